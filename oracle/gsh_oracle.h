/*
 * gsh_oracle.h -- CPU FP64 restatement of the SCHISM wave-equation hot path (test
 * infrastructure only: the parity checker and the CPU baseline, never the product path).
 *
 * Every function follows the reference's algorithm line by line; citations are in
 * gsh_oracle.c. Array layouts and field/array codes are the same as include/gsh.h.
 */
#ifndef GSH_ORACLE_H
#define GSH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_mesh or_mesh;

typedef struct or_params {
  double dt, theta, h0, ramp, dtb_min, dtb_max, velmin_btrack;
  int32_t ihhat, iettype;
  double rmaxvel;
  double dfv0;
  double time;
  double shapiro0;
  int32_t ishapiro;
} or_params;

const char* or_error(void);
/* Builds topology and geometry; NULL on failure (message in or_error()). kb may be NULL. */
or_mesh* or_mesh_new(int32_t n, int32_t ne, int32_t nvrt, const double* xy, const int32_t* elnode,
                     const double* depth, const int8_t* open_bnd, const int32_t* kb, const double* sigma);
void or_mesh_free(or_mesh* m);
/* Vertical coordinates beyond one sigma per level (readVGrid, read_and_process_grid.cpp:1313-1520):
 * ivcor 1 with sigma_node [n][nvrt] (mesh node order), or ivcor 2 (SZ) with nz Z levels ztot [nz],
 * S levels sigma_s [nvrt-nz+1] and h_s, h_c, theta_b, theta_f; for ivcor 2 every node's kb is
 * recomputed from its depth. Returns 0, or -1 with or_error(). */
int or_set_vertical(or_mesh* m, int32_t ivcor, const double* sigma_node, int32_t nz, double h_s, double h_c,
                    double theta_b, double theta_f, const double* ztot, const double* sigma_s);
int32_t or_mesh_ns(const or_mesh* m);
int64_t or_mesh_nnz(const or_mesh* m);
int32_t or_mesh_nbr_count(const or_mesh* m);
/* Copies a topology array (GSH_T_* codes) out. Returns 0 on success. */
int or_mesh_get(const or_mesh* m, int what, void* out);
/* State fields (GSH_F_* codes and layouts). Returns 0 on success. */
int or_set_field(or_mesh* m, int field, const void* data);
int or_get_field(const or_mesh* m, int field, void* out);

void or_levels(or_mesh* m, const or_params* p);
/* Backtracks every side x level (list == NULL) or the given (side, level) pairs. */
void or_backtrack(or_mesh* m, const or_params* p, const int32_t* list_side, const int32_t* list_lev,
                  int64_t count, int nthreads);
void or_rhs(or_mesh* m, const or_params* p);
/* tidal constituents of the open boundary (amig, ff, face [nbfr]; emo, efa [n][nbfr]) */
int or_set_tides(or_mesh* m, int32_t nbfr, const double* amig, const double* ff, const double* face,
                 const double* emo, const double* efa);
/* Nodal velocities (field 2) from side velocities (field 3) and element w (field 20), ball order. */
void or_nodal_velocity(or_mesh* m, const or_params* p, int nthreads);
/* side velocities u2 of the next step: elevation gradients, momentum solve, land-boundary slip */
void or_momentum(or_mesh* m, const or_params* p);
/* element vertical velocity w (GSH_F_ELEM_W) from the side velocities and the levels */
void or_vertical_velocity(or_mesh* m);
/* Assembly + essential BC into val (CSR of the mesh) and b. */
void or_assemble(or_mesh* m, const or_params* p, double* val, double* b, int nthreads);

/* Geometry arrays: 0 area [ne], 1 gradphi [ne][3][2], 2 side centre [ns][2], 3 side length [ns],
 * 4 side normal [ns][2]. Returns 0 on success. */
int or_geometry(const or_mesh* m, int what, double* out);
/* Element::interpolateVelocities at (x, y, z) in element e with the current state (test hook). */
void or_interp(const or_mesh* m, int32_t e, double x, double y, double z, double h0, double out[3]);

/* Jacobi PCG with the reference's recurrences; dot products in the fixed tile/wave tree of the
 * device kernels (rows_per_thread as the device solver's). h0/h1/h2 may be NULL (zero guess).
 * Returns the iteration count; *rr_last = last checked r.r, *atol2 = (rtol ||b||)^2. */
int32_t or_pcg(int32_t n, const int32_t* indptr, const int32_t* indices, const double* val, const double* b,
               const double* h0, const double* h1, const double* h2, double rtol, int32_t maxiter,
               int32_t check_every, int32_t rows_per_thread, double* x_out, double* rr_last, double* atol2,
               int nthreads);
/* Greedy first-fit colouring in smallest-last order of the graph of a CSR pattern (diagonal
 * ignored; ties by smaller index). Returns the number of colours. */
int32_t or_colour(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* colour);
/* PCG with the multicolour SSOR(omega) preconditioner applied by Eisenstat's trick, the Jacobi
 * driver's stopping rule on ||b - A x||, dot products in the device tree over the colour-major
 * padded order. Returns the iteration count, -1 on a bad pattern (or_error()). */
int32_t or_pcg_ssor(int32_t n, const int32_t* indptr, const int32_t* indices, const double* val, const double* b,
                    const double* h0, const double* h1, const double* h2, double omega, double rtol,
                    int32_t maxiter, int32_t check_every, int32_t rows_per_thread, double* x_out, double* rr_last,
                    double* atol2, int32_t* ncolours, int nthreads);
/* The same solver with its colouring and layout built once (rows_per_thread fixes the padding). */
typedef struct or_ssor or_ssor;
or_ssor* or_ssor_new(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t rows_per_thread);
int32_t or_ssor_ncolours(const or_ssor* s);
int32_t or_ssor_solve(or_ssor* s, const double* val, const double* b, const double* h0, const double* h1,
                      const double* h2, double omega, double rtol, int32_t maxiter, int32_t check_every,
                      double* x_out, double* rr_last, double* atol2, int nthreads);
void or_ssor_free(or_ssor* s);
/* Jacobi PCG of M (a power of two <= 256) systems on one CSR pattern, member-interleaved (values
 * [nnz][M], b and x [n][M]), cold start, dot products in the library's block order; iters[M],
 * rr_last[M]. Returns 0, or -1 with or_error(). */
int or_pcg_block(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t M, const double* val,
                 const double* b, double rtol, int32_t maxiter, int32_t check_every, double* x, int32_t* iters,
                 double* rr_last, int nthreads);
/* Dot product in the device's fixed reduction order. */
double or_dot_canon(int32_t n, const double* a, const double* b, int32_t rows_per_thread);
/* ||b - A x|| / ||b|| in the same order as the device check. */
double or_true_rel(int32_t n, const int32_t* indptr, const int32_t* indices, const double* val, const double* b,
                   const double* x, int32_t rows_per_thread);

#ifdef __cplusplus
}
#endif
#endif
