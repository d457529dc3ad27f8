/*
 * A native caller of the C ABI (include/gsh.h): one time step of the wave-equation path on a
 * synthetic channel, every call's status checked, from plain C with no Python in the process.
 *
 *   step_harness --version        library version only (no device use)
 *   step_harness NX NY NVRT       build the mesh, run backtracking, RHS producers, assembly,
 *                                 the FP64 SSOR solve, the elevation update, momentum, element
 *                                 w, levels and nodal velocities; prints one summary line
 *
 * Exit status 0 on success, 1 on a failed call (its gsh_last_error is printed).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gsh.h"

#define CHECK(call)                                                                      \
  do {                                                                                   \
    int rc_ = (call);                                                                    \
    if (rc_ != GSH_OK) {                                                                 \
      fprintf(stderr, "%s failed with %d: %s\n", #call, rc_, gsh_last_error());          \
      exit(1);                                                                           \
    }                                                                                    \
  } while (0)

int main(int argc, char** argv) {
  if (argc == 2 && strcmp(argv[1], "--version") == 0) {
    printf("gsh %d\n", gsh_version());
    return 0;
  }
  if (argc != 4) {
    fprintf(stderr, "usage: %s --version | NX NY NVRT\n", argv[0]);
    return 2;
  }
  const int nx = atoi(argv[1]), ny = atoi(argv[2]), nvrt = atoi(argv[3]);
  if (nx < 4 || ny < 4 || nvrt < 2 || nvrt > 1024) {
    fprintf(stderr, "bad sizes\n");
    return 2;
  }
  const double dx = 500.0;

  /* synthetic channel: nodes, triangles, open west edge, bathymetry, sigma levels */
  gsh_gen* g = NULL;
  CHECK(gsh_gen_channel(nx, ny, dx, 0.2, 11, &g));
  int32_t n = 0, ne = 0;
  CHECK(gsh_gen_sizes(g, &n, &ne));
  double* xy = malloc(sizeof(double) * 2 * n);
  int32_t* en = malloc(sizeof(int32_t) * 3 * ne);
  int8_t* ob = malloc(n);
  CHECK(gsh_gen_get(g, xy, en, ob));
  CHECK(gsh_gen_destroy(g));
  double* depth = malloc(sizeof(double) * n);
  CHECK(gsh_gen_bathymetry(nx, ny, 20.0, 5.0, 50.0, 3, depth));
  double* sigma = malloc(sizeof(double) * nvrt);
  for (int k = 0; k < nvrt; ++k) sigma[k] = -1.0 + (double)k / (nvrt - 1);
  sigma[nvrt - 1] = 0.0;

  gsh_ctx* ctx = NULL;
  CHECK(gsh_ctx_create(0, &ctx));
  gsh_mesh_desc d;
  memset(&d, 0, sizeof d);
  d.n = n;
  d.ne = ne;
  d.nvrt = nvrt;
  d.xy = xy;
  d.elnode = en;
  d.depth = depth;
  d.open_bnd = ob;
  d.sigma = sigma;
  d.kb = NULL;
  d.renumber = 1;
  gsh_mesh* mesh = NULL;
  int32_t* perm = malloc(sizeof(int32_t) * n);
  int32_t* eperm = malloc(sizeof(int32_t) * ne);
  CHECK(gsh_mesh_create(ctx, &d, &mesh, perm, eperm));
  gsh_mesh_info info;
  CHECK(gsh_mesh_get_info(mesh, &info));

  /* state: a small elevation wave and a uniform along-channel flow at every side and node */
  double* eta = malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) eta[i] = 0.05 * sin(xy[2 * perm[i]] / (8.0 * dx));
  CHECK(gsh_mesh_set_field(mesh, GSH_F_ETA2, eta, n));
  CHECK(gsh_mesh_set_field(mesh, GSH_F_ETA1, eta, n));
  const int64_t nsv = (int64_t)info.ns * nvrt, nnv = (int64_t)n * nvrt;
  double* su2 = malloc(sizeof(double) * 2 * nsv);
  for (int64_t j = 0; j < nsv; ++j) su2[2 * j] = 0.3, su2[2 * j + 1] = 0.02;
  CHECK(gsh_mesh_set_field(mesh, GSH_F_SIDE_U2, su2, 2 * nsv));
  double* uvw = malloc(sizeof(double) * 3 * nnv);
  for (int64_t j = 0; j < nnv; ++j) uvw[3 * j] = 0.3, uvw[3 * j + 1] = 0.02, uvw[3 * j + 2] = 0.0;
  CHECK(gsh_mesh_set_field(mesh, GSH_F_NODE_UVW, uvw, 3 * nnv));

  gsh_step_params p;
  gsh_step_params_default(&p);
  CHECK(gsh_update_levels(mesh, &p));
  gsh_solver* s = NULL;
  CHECK(gsh_solver_create(mesh, GSH_PREC_FP64, GSH_PRECOND_SSOR, &s));

  /* one step in schism_step order */
  CHECK(gsh_backtrack(mesh, &p, NULL, NULL));
  CHECK(gsh_rhs_producers(mesh, &p));
  CHECK(gsh_assemble(mesh, s, &p, 0.0));
  int iters = 0, status = 0;
  CHECK(gsh_solver_solve(s, NULL, NULL, 1e-12, 0, 10, 60.0, &iters, &status));
  double rel = 0.0;
  CHECK(gsh_solver_check(s, &rel));
  CHECK(gsh_update_elevation(mesh, s));
  CHECK(gsh_momentum(mesh, &p));
  CHECK(gsh_vertical_velocity(mesh, &p));
  CHECK(gsh_update_levels(mesh, &p));
  CHECK(gsh_nodal_velocity(mesh, &p));
  CHECK(gsh_ctx_synchronize(ctx));

  /* a few sanity figures of the new state */
  CHECK(gsh_mesh_get_field(mesh, GSH_F_SIDE_U2, su2, 2 * nsv));
  double umax = 0.0;
  int finite = 1;
  for (int64_t j = 0; j < 2 * nsv; ++j) {
    if (!isfinite(su2[j])) finite = 0;
    if (fabs(su2[j]) > umax) umax = fabs(su2[j]);
  }
  printf("n=%d ne=%d ns=%d nvrt=%d iters=%d status=%d rel=%.3e max|u2|=%.4f finite=%d\n", n, ne, info.ns, nvrt,
         iters, status, rel, umax, finite);

  CHECK(gsh_solver_destroy(s));
  CHECK(gsh_mesh_destroy(mesh));
  CHECK(gsh_ctx_destroy(ctx));
  free(xy);
  free(en);
  free(ob);
  free(depth);
  free(sigma);
  free(perm);
  free(eperm);
  free(eta);
  free(su2);
  free(uvw);
  return (finite && status == 0 && rel < 1e-10) ? 0 : 1;
}
