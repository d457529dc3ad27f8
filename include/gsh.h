/*
 * gsh.h -- C ABI of the MI355X-native SCHISM wave-equation hot path.
 *
 * One per-timestep path of SCHISM/gschism is rebuilt here:
 *   - ELM backtracking from side centres     (reference: src/kernels/meshdata.cu:3547-3660, 1449-1629)
 *   - RHS producers U, chi/H^, G^            (meshdata.cu:4459-4476, 4615-4725, 6183-6424)
 *   - wave-equation assembly + essential BC  (meshdata.cu:3182-3474, 4276-4339, 4400-4446)
 *   - Jacobi-preconditioned CG solve         (src/solver/pcg_solver.py:1075-1124, pcg_loop.cpp:334-416)
 *   - elevation update                       (meshdata.cu:3538-3544, 4449-4456)
 *
 * Conventions
 *   - Every entry point returns GSH_OK (0) or a GSH_E_* status; the message of the last
 *     failure on the calling thread is returned by gsh_last_error(). No C++ exception crosses
 *     this boundary.
 *   - Plain pointers and sizes only. Pointers named *_host are host memory; pointers returned
 *     by *_ptr / gsh_solver_csr are device (HBM) addresses on the context's GPU.
 *   - Node, element and side indices are 0-based. Per-level arrays are stored entity-major,
 *     level-minor ([entity][level]) as in the reference's Node/Side structs
 *     (meshdata.h:1577-1580, 1324-1325).
 *   - A handle is single-stream and not re-entrant; distinct handles may be used from
 *     different threads.
 */
#ifndef GSH_H
#define GSH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define GSH_OK            0
#define GSH_E_ARG         1   /* invalid argument (reference: ValueError / RuntimeError on bad input) */
#define GSH_E_DEVICE      2   /* HIP runtime failure (reference: cuCheck, pcg_loop.cpp:46-54) */
#define GSH_E_OOM         3   /* device allocation failed */
#define GSH_E_TIMEOUT     4   /* GPU did not finish in time (reference: waitStream, pcg_loop.cpp:444-456) */
#define GSH_E_ADMISSION   5   /* serving: request does not fit the device memory budget */
#define GSH_E_TOPOLOGY    6   /* broken ball / hanging node / orientation (read_and_process_grid.cpp:47-58, 154-159, 334-341) */
#define GSH_E_STATE       7   /* call out of order (e.g. solve before a matrix exists) */
#define GSH_E_NOTFOUND    8   /* serving: unknown ticket */
#define GSH_E_IO          9   /* a grid file cannot be read or breaks its format (reference: runtime_error
                                 in readHGrid / readVGrid, read_and_process_grid.cpp:864, 963-966, 1050-1053) */
#define GSH_W_MAXITER     64  /* flag OR-ed into *status_out of a solve that stopped at maxiter
                                 (the reference returns maxiter silently, pcg_solver.py:1068) */

#define GSH_PREC_FP64     0   /* plain FP64 PCG */
#define GSH_PREC_MIXED    1   /* FP32 SpMV/precond/axpy, FP64 dots + true residual + refinement */

#define GSH_PRECOND_JACOBI 0  /* reference preconditioner (pcg_solver.py:933-938) */
#define GSH_PRECOND_SSOR   1  /* multicolour symmetric SOR; FP64, or FP32 inner solves under GSH_PREC_MIXED
                                 (single systems; no counterpart in the reference,
                                 whose only preconditioner is Jacobi; same driver and stopping rule) */

/* ---------------------------------------------------------------- context */
typedef struct gsh_ctx gsh_ctx;
typedef struct gsh_mesh gsh_mesh;
typedef struct gsh_solver gsh_solver;

int  gsh_ctx_create(int device, gsh_ctx** out);
int  gsh_ctx_destroy(gsh_ctx* ctx);
/* Message of the last failure on this thread ("" if none). */
const char* gsh_last_error(void);
/* Library version as 10000*major + 100*minor + patch. */
int  gsh_version(void);
/* Device memory figures of the context's GPU, in bytes. */
int  gsh_ctx_mem_info(gsh_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);
/* Waits until all work issued on the context's stream has finished. */
int  gsh_ctx_synchronize(gsh_ctx* ctx);

/* ---------------------------------------------------------------- mesh */
/* Host description of a triangular mesh. Replaces the hgrid/vgrid arrays that
 * readHGrid/readVGrid hand to populateNodes/populateSides/populateElements
 * (read_and_process_grid.cpp:829-1249, meshdata.cu:1903-2462). */
/* Vertical grid beyond one sigma per level, as readVGrid returns it
 * (read_and_process_grid.cpp:1313-1520) and Node::compZCoord uses it (meshdata.cu:1744-1807). */
typedef struct gsh_vgrid {
  int32_t ivcor;             /* 1 = localized sigma (LSC2, a sigma per node and level), 2 = SZ */
  /* ivcor = 1 */
  const double* sigma_node;  /* [n][nvrt] sigma of every node and level (desc node order);
                                levels at or below a node's kb are its bed */
  /* ivcor = 2: levels 0 .. nz-2 are Z levels ztot, level nz-1 is at -min(h, h_s), levels
   * nz .. nvrt-1 are S levels 1 .. nvrt-nz of sigma_s; each node's kb follows from its depth
   * (Node::updateKb, meshdata.cu:1676-1718) and desc->kb is ignored */
  int32_t nz;                /* kz of vgrid.in: Z levels including the one at -h_s (>= 1) */
  double h_s, h_c, theta_b, theta_f;
  const double* ztot;        /* [nz] increasing, ztot[k] < -h_s for k < nz-1; ztot[nz-1] = -h_s */
  const double* sigma_s;     /* [nvrt-nz+1] S levels, sigma_s[0] = -1 ... sigma_s[nvrt-nz] = 0 */
} gsh_vgrid;

typedef struct gsh_mesh_desc {
  int32_t n;              /* nodes */
  int32_t ne;             /* triangles */
  int32_t nvrt;           /* vertical levels (2 .. 1024) */
  const double*  xy;      /* [n][2] node coordinates (ics=1 Cartesian, metres) */
  const int32_t* elnode;  /* [ne][3] counter-clockwise node ids */
  const double*  depth;   /* [n] depth h (positive down, metres) */
  const int8_t*  open_bnd;/* [n] 1 = node on an open boundary with an elevation (essential) BC; NULL = none */
  const double*  sigma;   /* [nvrt] sigma levels, sigma[0]=-1 ... sigma[nvrt-1]=0 (ivcor=1, meshdata.cu:1798-1803);
                             may be NULL when vgrid is given */
  const int32_t* kb;      /* [n] bottom level index per node; NULL = 0 everywhere */
  int32_t renumber;       /* 1 = reverse Cuthill-McKee renumbering of nodes (elements follow) */
  const gsh_vgrid* vgrid; /* NULL = ivcor 1 with `sigma` at every node */
} gsh_mesh_desc;

typedef struct gsh_mesh_info {
  int32_t n, ne, ns, nvrt;
  int64_t nnz;            /* CSR nonzeros = sum(numNodes+1) (main.py:262-279) */
  int32_t max_row;        /* longest CSR row */
  int32_t n_open;         /* nodes with an essential BC */
} gsh_mesh_info;

/* Builds the mesh on the device: topology (CCW balls, sides, CSR pattern) exactly as
 * read_and_process_grid.cpp:27-174, 272-437, 496-574 and meshdata.cu:2094-2141 define it,
 * element/side geometry (meshdata.cu:2274-2321, 2350-2462). With desc->renumber the node
 * order is RCM; node_perm_out[new] = old node id and elem_perm_out[new] = old element id
 * (either may be NULL). */
int gsh_mesh_create(gsh_ctx* ctx, const gsh_mesh_desc* desc, gsh_mesh** out,
                    int32_t* node_perm_out, int32_t* elem_perm_out);
int gsh_mesh_destroy(gsh_mesh* mesh);
int gsh_mesh_get_info(const gsh_mesh* mesh, gsh_mesh_info* out);

/* Topology arrays (host copies), selected by GSH_T_*; sizes from gsh_mesh_get_info. */
#define GSH_T_XY           0   /* double [n][2] (renumbered)                          */
#define GSH_T_ELNODE       1   /* int32  [ne][3]                                       */
#define GSH_T_ELSIDE       2   /* int32  [ne][3]  side j is opposite node j            */
#define GSH_T_ISDEL        3   /* int32  [ns][2]  elements of a side, -1 = none        */
#define GSH_T_ISIDENODE    4   /* int32  [ns][2]                                       */
#define GSH_T_BALL_PTR     5   /* int32  [n+1]    CCW ball of elements                 */
#define GSH_T_BALL_ELEM    6   /* int32  [ball_ptr[n]]                                 */
#define GSH_T_NBR_NODE     7   /* int32  [ball-node count]  (row i spans indptr[i]-i .. indptr[i+1]-i-1) */
#define GSH_T_CSR_INDPTR   8   /* int32  [n+1]                                         */
#define GSH_T_CSR_INDICES  9   /* int32  [nnz]                                         */
#define GSH_T_CSR_SLOT    10   /* int16  [nnz-n]  slot of neighbour j in its CSR row   */
#define GSH_T_DEPTH       11   /* double [n] (renumbered)                              */
#define GSH_T_OPEN_BND    12   /* int8   [n]                                           */
#define GSH_T_KB          13   /* int32  [n]                                           */
#define GSH_T_SIDE_TYPE   14   /* int8   [ns] 0 internal, 1 land boundary, 2 open boundary */
int gsh_mesh_get_topology(const gsh_mesh* mesh, int what, void* host_out);

/* Per-entity state fields (device resident, host copies in and out). */
#define GSH_F_ETA2        0   /* double [n]          elevation at the current step     */
#define GSH_F_ETA1        1   /* double [n]          elevation at the previous step    */
#define GSH_F_NODE_UVW    2   /* double [n][nvrt][3] nodal u2.x, u2.y, w2              */
#define GSH_F_SIDE_U2     3   /* double [ns][nvrt][2] side velocity u2                 */
#define GSH_F_BTU         4   /* double [ns][nvrt][2] backtracked velocity (output)    */
#define GSH_F_FOOT        5   /* int32  [ns][nvrt]   foot element of each backtrack, -1 = not tracked */
#define GSH_F_BSTATUS     6   /* int32  [ns][nvrt]   backtrack flags GSH_BT_*          */
#define GSH_F_CD          7   /* double [n]          bottom drag coefficient (nchi=0)  */
#define GSH_F_GRAV2       8   /* double [n]          gravity                           */
#define GSH_F_ETH         9   /* double [n]          essential-BC elevation target     */
#define GSH_F_VSOURCE    10   /* double [ne]         volumetric source                 */
#define GSH_F_SIDE_U     11   /* double [ns][2]      vertically integrated velocity    */
#define GSH_F_SIDE_HHAT  12   /* double [ns]                                           */
#define GSH_F_SIDE_CHI   13   /* double [ns]                                           */
#define GSH_F_GHAT1      14   /* double [ne][2]                                        */
#define GSH_F_QEL        15   /* double [n]          assembled RHS                     */
#define GSH_F_ZC         16   /* double [ns][nvrt]   side-centre z                     */
#define GSH_F_ELBC       17   /* double [n]          computed boundary elevation       */
#define GSH_F_WET        18   /* int8   [n]/[ne]/[ns] packed: nodes, then elements, then sides (1 = wet) */
#define GSH_F_DEPTH      19   /* double [n]          still-water depth h (dp); levels and wet flags follow at
                                 the next gsh_update_levels */
#define GSH_F_ELEM_W     20   /* double [ne][nvrt]   element vertical velocity w (gsh_vertical_velocity, or set
                                 by the caller); allocated at first use, zero until written */
#define GSH_F_ETYPE      21   /* int8   [n]          elevation BC type of each open node: 0 none, 1-5 as
                                                       iettype of bctides.in (readTides.py:161-185);
                                                       -1 (the initial value) = gsh_step_params.iettype */
int gsh_mesh_set_field(gsh_mesh* mesh, int field, const void* host, int64_t count);
int gsh_mesh_get_field(const gsh_mesh* mesh, int field, void* host, int64_t count);
int gsh_mesh_field_ptr(gsh_mesh* mesh, int field, void** dev_ptr, int64_t* count);

/* Backtrack status bits (GSH_F_BSTATUS) for paths the reference only asserts on. */
#define GSH_BT_WALK_CAP   1   /* celestial walk hit its 40-step cap (meshdata.cu:87,126)   */
#define GSH_BT_ITER38     2   /* btrack returned at iter==38 (meshdata.cu:1611)           */
#define GSH_BT_STUCK      4   /* time_rm stopped decreasing (meshdata.cu:1496-1500)        */
#define GSH_BT_OPEN_EXIT  8   /* track left the domain through an open boundary (:1567)    */
#define GSH_BT_SLIDE     16   /* track slid along a land boundary or dry edge (:1569-1584) */

/* Physical parameters of one step (the constexpr values of constants_template.h:140-168). */
typedef struct gsh_step_params {
  double dt;          /* time step [s]                                 */
  double theta;       /* implicitness thetai                           */
  double h0;          /* wet/dry threshold [m]                         */
  double ramp;        /* boundary ramp factor (1 = none)               */
  double dtb_min;     /* ELM minimum sub-step [s]                      */
  double dtb_max;     /* ELM maximum sub-step [s]                      */
  double velmin_btrack; /* ELM minimum velocity [m/s]                  */
  int32_t ihhat;      /* 1 = clip H^ >= 0 in 3D                        */
  int32_t iettype;    /* elevation BC type of the open boundary: 1,2,4 = eth; 0 = none */
  double rmaxvel;     /* velocity clamp of the nodal reconstruction: |u| <= rmaxvel,
                         |v| <= 1.013 rmaxvel (constants_template.h:142-144)     */
  double dfv0;        /* vertical viscosity of the momentum solve, itur=0 [m^2/s]  */
  double time;        /* model time [s] of the harmonic elevation BC (iettype 3, 5) */
  double shapiro0;    /* Shapiro filter strength of the side velocities (readNamelist.py:95) */
  int32_t ishapiro;   /* 1 = gsh_momentum applies the Shapiro filter (readNamelist.py:94)  */
} gsh_step_params;
void gsh_step_params_default(gsh_step_params* p);

/* a3: node z-levels, wet/dry of nodes, elements and sides, side kb and zc
 * (meshdata.cu:5521-5594, 5679-5760, 1744-1807). Call after eta2 changes. */
int gsh_update_levels(gsh_mesh* mesh, const gsh_step_params* p);

/* a4-a7: backtrackAllNodes (meshdata.cu:3547-3660) over every side x level. Results stay on
 * the device (GSH_F_BTU, GSH_F_FOOT, GSH_F_BSTATUS); btu_host / foot_host may be NULL. */
int gsh_backtrack(gsh_mesh* mesh, const gsh_step_params* p, double* btu_host, int32_t* foot_host);

/* How gsh_backtrack reads the elements: 0 (default) through the caches; 1 stages each
 * workgroup's element patch (the elements on both sides of its sides and their neighbours) in
 * LDS first and reads the rest from global memory. Results are identical; only speed differs.
 * No counterpart in the reference (a tuning choice of this implementation). */
int gsh_mesh_set_backtrack_staging(gsh_mesh* mesh, int32_t staged);

/* Backtracks an explicit list of (side, level) start points (the same walk as gsh_backtrack,
 * results written to the same per-(side, level) slots). */
int gsh_backtrack_list(gsh_mesh* mesh, const gsh_step_params* p, const int32_t* sides_host,
                       const int32_t* levels_host, int32_t count);

/* a8-a10 and the elevation BC: computeVerticallyIntegratedVelocity, computeVegetationAndDragParameters,
 * computeElemAverageGhat and computeElevationBoundaryConditions (meshdata.cu:4459-4476, 4615-4725,
 * 6183-6424, 4276-4339). */
int gsh_rhs_producers(gsh_mesh* mesh, const gsh_step_params* p);

/* Tidal forcing of the open boundary for iettype 3 and 5 (bctides.in, readTides.py:96-118,
 * 170-185; used by computeElevationBoundaryConditions, meshdata.cu:4298-4316): nbfr
 * constituents with angular frequency amig, nodal factor ff and equilibrium argument face
 * [rad], and per node amplitude emo [n][nbfr] and phase efa [n][nbfr] [rad] (read at open
 * nodes only). nbfr = 0 removes the forcing; nbfr <= 256. */
int gsh_set_tides(gsh_mesh* mesh, int32_t nbfr, const double* amig, const double* ff, const double* face,
                  const double* emo, const double* efa);

/* The step before backtracking: nodal velocities from side and element velocities.
 * zeroNodalVelocities, computeNodalVerticalVelocities1/2 and computeNodalHorizontalVelocities1/2
 * (meshdata.cu:5765-5929) for triangles, ics=1, indvel <= 0: GSH_F_NODE_UVW from GSH_F_SIDE_U2 and
 * GSH_F_ELEM_W, with the wet flags and levels of the last gsh_update_levels. The reference adds the
 * element contributions with atomics; here each node sums its ball elements in ball order. */
int gsh_nodal_velocity(gsh_mesh* mesh, const gsh_step_params* p);

/* The step after the solve (call after gsh_update_elevation, before gsh_update_levels):
 * computeHorGradientAtSide, solveMomentumEquations and applyMomentumBC (meshdata.cu:5045-5446)
 * for the barotropic, itur=0 (dfv0), iveg=0, ics=1, imm=0, islip=0 subset without Coriolis,
 * wind, horizontal viscosity, pressure or tidal potential: GSH_F_SIDE_U2 from eta1/eta2, btu,
 * chi and H^ (gsh_rhs_producers) and the levels and wet flags of the last gsh_update_levels.
 * 3D sides solve their vertical tridiagonal system, 2D sides the depth-averaged update; land
 * boundary sides keep the tangential component. */
int gsh_momentum(gsh_mesh* mesh, const gsh_step_params* p);
/* With p->ishapiro == 1, gsh_momentum ends with shapiroFilter1/2 (meshdata.cu:5456-5518, called at
 * schism_step.py:289-292): every wet internal side's u2[k], k > kb, becomes
 * u2 + shapiro0 (sum of its four neighbour sides' u2 / 4 - u2), the neighbours' levels clamped to
 * max(k, kb + 1) where wet; the 2D bottom slot follows the surface value. The filtered field is
 * written into a second buffer that then becomes GSH_F_SIDE_U2, so a device pointer to that field
 * (gsh_mesh_field_ptr) must be fetched again after the call. */

/* computeVerticalVelocity (meshdata.cu:5978-6102), triangles, ics=1, imm=0: GSH_F_ELEM_W from
 * GSH_F_SIDE_U2 (after gsh_momentum) and the levels of the last gsh_update_levels (node z from
 * the elevation of that update, side zc and kb): the flux balance of each prism bottom up, with
 * w = -grad(depth) . u at the element's bottom level. */
int gsh_vertical_velocity(gsh_mesh* mesh, const gsh_step_params* p);

/* ---------------------------------------------------------------- grid files (host side)
 * hgrid.gr3 as readHGrid reads it (read_and_process_grid.cpp:855-1070): node coordinates and
 * depth, triangles (1-based node ids in the file, 0-based here), open and land boundary segments,
 * with its checks (neta and nvel totals, looped open boundaries); quadrilaterals are rejected.
 * vgrid.in with ivcor=1 as readVGrid reads it (:1337-1385): kbp per node and sigma rows, accepted
 * when the nodes agree on sigma at every level they use (the library holds one sigma per level
 * and a bottom level per node); ivcor=2 is rejected. vgrid_path may be NULL. */
typedef struct gsh_grid gsh_grid;
int gsh_grid_read(const char* hgrid_path, const char* vgrid_path, gsh_grid** out);
int gsh_grid_info(const gsh_grid* g, int32_t* n, int32_t* ne, int32_t* nvrt, int32_t* n_open_seg,
                  int32_t* n_land_seg);
/* Any output may be NULL; kb [n] and sigma [nvrt] need a vgrid. open_bnd[i] = 1 on open segments. */
/* kb is in the file for ivcor 1 only; sigma (one per level) exists for ivcor 1 when every node has
 * the same column above its bed (GSH_E_STATE otherwise). */
int gsh_grid_get(const gsh_grid* g, double* xy, double* depth, int32_t* elnode, int8_t* open_bnd, int32_t* kb,
                 double* sigma);
/* Nodes (0-based) of open (land = 0) or land (land = 1) segment seg; *island = the land flag. */
/* The vertical grid of the vgrid.in read (readVGrid, read_and_process_grid.cpp:1313-1520): ivcor
 * (1 or 2); ivcor 1: per_node = 1 when the nodes' sigma differ above their beds, and then
 * sigma_node [n][nvrt] receives every node's column (gsh_grid_get's kb and, with per_node = 0,
 * sigma give the rest); ivcor 2: nz (kz), params = {h_s, h_c, theta_b, theta_f}, ztot [nz],
 * sigma_s [nvrt-nz+1]. Outputs may be NULL. */
int gsh_grid_vgrid(const gsh_grid* g, int32_t* ivcor, int32_t* per_node, int32_t* nz, double* params,
                   double* ztot, double* sigma_s, double* sigma_node);
int gsh_grid_segment(const gsh_grid* g, int land, int32_t seg, int32_t* nodes_out, int32_t cap, int32_t* count,
                     int32_t* island);
int gsh_grid_destroy(gsh_grid* g);

/* ---------------------------------------------------------------- solver */
/* PCGSolverMultiGPU(configCuda, N, NNZ) equivalent bound to a mesh's CSR pattern
 * (pcg_solver.py:547-671). */
int gsh_solver_create(gsh_mesh* mesh, int precision, int precond, gsh_solver** out);
/* Stand-alone solver over a caller CSR pattern (int32 indptr/indices, host arrays); values
 * are set with gsh_solver_set_values or written through gsh_solver_csr. */
int gsh_solver_create_csr(gsh_ctx* ctx, int32_t n, int64_t nnz, const int32_t* indptr_host,
                          const int32_t* indices_host, int precision, int precond, gsh_solver** out);
int gsh_solver_destroy(gsh_solver* s);
/* csr_data / csr_indices / csr_indptr (pcg_solver.py:1146-1148): device pointers. */
int gsh_solver_csr(gsh_solver* s, double** data, int32_t** indices, int32_t** indptr, int64_t* nnz);
int gsh_solver_set_values(gsh_solver* s, const double* data_host);
int gsh_solver_get_values(gsh_solver* s, double* data_host);
/* Right-hand side b and solution x as device vectors (length n). */
int gsh_solver_vec_ptr(gsh_solver* s, double** b, double** x);
/* Host copy of the right-hand side b (gatherBFromMesh target, pcg_solver.py:950-954). */
int gsh_solver_get_rhs(gsh_solver* s, double* b_host);
/* Sets b from the host (PCGSolverMultiGPU.setRHS of the mirror; the next solve uses it). */
int gsh_solver_set_rhs(gsh_solver* s, const double* b_host);
/* Host copy of the last solution (scatterXToMesh source, pcg_solver.py:957-961). */
int gsh_solver_get_solution(gsh_solver* s, double* x_host);

/* a11+a12: setupCoefficientMatrixAndRHS + setEssentialBC (meshdata.cu:3182-3474, 4400-4446)
 * into the solver's CSR values and b (= qel). */
int gsh_assemble(gsh_mesh* mesh, gsh_solver* s, const gsh_step_params* p, double time);

/* solve(rtol, maxiter, checkEvery) (pcg_solver.py:1075-1124; native loop pcg_loop.cpp:253-291):
 * b from b_host if non-NULL, else the device b (the assembled qel, gatherBFromMesh);
 * initial guess 3x0-3x1+x2 from the solver's history; returns the iteration count in
 * *iters_out (maxiter if not converged, GSH_W_MAXITER set in *status_out). x_host may be NULL.
 * timeout_s <= 0 means 60 s. */
int gsh_solver_solve(gsh_solver* s, const double* b_host, double* x_host, double rtol,
                     int maxiter, int check_every, double timeout_s, int* iters_out, int* status_out);
/* No counterpart in the reference: a solve that allocates nothing, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). It gives the same iteration count, status and solution, bit for bit, as
 * gsh_solver_solve of an FP64 Jacobi solver made by gsh_solver_create_csr on this pattern with
 * its history cleared (zero initial guess) and these values and b. indptr (n+1), indices (nnz),
 * data (nnz), b (n) and x (n, the output) are DEVICE pointers on ctx's device; the pattern is
 * checked on the device (GSH_E_ARG). Every other buffer of the solve lives in one static device
 * array of the library (about 53 MB per device, held until the process ends) sized for
 * n <= GSH_STATIC_MAX_ROWS (GSH_E_ARG beyond). All calls on one device share that array: calls
 * that overlap on one device (several contexts or threads) race; consecutive calls do not. */
#define GSH_STATIC_MAX_ROWS 1048576
int gsh_pcg_solve_static(gsh_ctx* ctx, int32_t n, int64_t nnz, const int32_t* indptr_dev, const int32_t* indices_dev,
                         const double* data_dev, const double* b_dev, double* x_dev, double rtol, int maxiter,
                         int check_every, double timeout_s, int* iters_out, int* status_out);
/* checkSolution (pcg_solver.py:1127-1143): ||b - A x|| / ||b||. */
int gsh_solver_check(gsh_solver* s, double* rel_out);
/* Clears the extrapolation history x0=x1=x2=0 (cold start, main.py:352). */
int gsh_solver_reset_history(gsh_solver* s);
/* Per-solve statistics of the last solve. */
typedef struct gsh_solve_stats {
  int32_t iters;          /* PCG iterations (inner iterations for MIXED)      */
  int32_t outer;          /* refinement steps (MIXED), 1 for FP64             */
  double  rr;             /* final recursive ||r||^2                          */
  double  atol2;          /* (rtol ||b||)^2                                   */
  double  true_rel;       /* ||b-Ax||/||b|| if measured, else -1              */
  double  ms;             /* wall time of the solve                           */
} gsh_solve_stats;
int gsh_solver_last_stats(gsh_solver* s, gsh_solve_stats* out);
/* Tuning knobs: rows per thread of the PCG tiles (1, 2, 4, 8) and the inner tolerance of the
 * mixed-precision refinement. */
int gsh_solver_set_tuning(gsh_solver* s, int rows_per_thread, double inner_rtol);
/* Autotuner (no counterpart in the reference): times `iters` FP64 iterations on the solver's
 * current system for rows_per_thread 1, 2, 4 and 8 (kernel-stamped events) and keeps the fastest;
 * us_per_iter[4] (may be NULL) receives the measured microseconds per iteration. The solution
 * buffer is overwritten, the history of previous solutions is kept. */
int gsh_solver_autotune(gsh_solver* s, int iters, int* best_rpt, double* us_per_iter);
/* No counterpart in the reference: on != 0 captures each solve window (check_every iterations
 * and the window-end test) of this solver into a HIP graph once and replays it; default off. */
int gsh_solver_set_graphs(gsh_solver* s, int on);
/* SSOR relaxation factor omega in (0, 2), default 1 (symmetric Gauss-Seidel). */
int gsh_solver_set_omega(gsh_solver* s, double omega);
/* SSOR layout: number of colours, rows after padding every colour block to whole reduction
 * tiles, and the colour of every row (n entries, may be NULL). */
int gsh_solver_ssor_info(gsh_solver* s, int32_t* ncolours, int32_t* padded_rows, int32_t* colour_out);
/* Host only: the SSOR solver's colouring of a CSR pattern (greedy first fit in smallest-last
 * order, ties by smaller index; the diagonal is ignored). */
int gsh_colour_csr(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* colour_out,
                   int32_t* ncolours);

/* updateElevation (meshdata.cu:4449-4456) after rollUpElevation (:3538-3544):
 * eta1 <- eta2, eta2 <- x. */
int gsh_update_elevation(gsh_mesh* mesh, gsh_solver* s);

/* ---------------------------------------------------------------- batching */
/* Solves several independent systems in lock-step launches: every PCG phase (Jacobi) or colour
 * step of the SSOR sweeps is one launch over the tiles of all systems that have not converged.
 * Each system keeps the single-solve stopping rule, so results equal solving them one by one.
 * All solvers share one context, precision and preconditioner; SSOR solvers also share the rows
 * per thread. Systems with fewer colours sit out the colour steps beyond their own. */
int gsh_batch_solve(gsh_solver* const* solvers, int count, double rtol, int maxiter,
                    int check_every, double timeout_s, int* iters_out);

/* ---------------------------------------------------------------- multi-RHS block */
/* No counterpart in the reference (it solves one system per call): `members` systems (1-256) on
 * one mesh's CSR pattern, e.g. ensemble members with per-member bathymetry and right-hand side,
 * solved together by Jacobi PCG with the reference's recurrences, cold start and stopping rule per
 * member (pcg_solver.py:881-947, 1075-1124), in a member-interleaved layout that reads the pattern
 * once for all members. Dot products are summed in the block's own fixed order, so each member's
 * result equals the oracle's block restatement bit for bit and a single solve of that system to
 * rounding. */
typedef struct gsh_block gsh_block;
int gsh_block_create(gsh_mesh* mesh, int32_t members, gsh_block** out);
int gsh_block_destroy(gsh_block* b);
/* setupCoefficientMatrixAndRHS + setEssentialBC (meshdata.cu:3182-3474, 4400-4446) of the mesh's
 * current state into member `member`. */
int gsh_block_assemble(gsh_block* b, int32_t member, const gsh_step_params* p, double time);
/* A member's CSR values (nnz, the mesh's pattern) and/or right-hand side (n) from the host; NULL
 * keeps the current one. */
int gsh_block_set_member(gsh_block* b, int32_t member, const double* values_host, const double* b_host);
/* Solves every member from x = 0: iters_out[m] (maxiter if not converged) and status_out[m]
 * (GSH_W_MAXITER then); either may be NULL. timeout_s <= 0 means 60 s. */
int gsh_block_solve(gsh_block* b, double rtol, int maxiter, int check_every, double timeout_s,
                    int32_t* iters_out, int32_t* status_out);
int gsh_block_get_solution(gsh_block* b, int32_t member, double* x_host);

/* ---------------------------------------------------------------- serving */
typedef struct gsh_server gsh_server;
typedef struct gsh_server_config {
  int64_t hbm_budget_bytes;   /* admission budget (0 = free device memory at creation) */
  int32_t max_active;         /* max systems in flight (0 = unlimited)             */
  int32_t precision;          /* GSH_PREC_*                                         */
  double  rtol;               /* solve tolerance                                    */
  int32_t maxiter;
  int32_t check_every;
  int32_t precond;            /* GSH_PRECOND_* (Jacobi: 4 rows per thread, SSOR: 2)   */
} gsh_server_config;
int gsh_server_create(gsh_ctx* ctx, const gsh_server_config* cfg, gsh_server** out);
int gsh_server_destroy(gsh_server* srv);
/* Registers a mesh so that requests can reference it; the mesh stays device resident. Its
 * requests' set-up (backtracking, RHS producers, assembly) uses gsh_step_params_default until
 * gsh_server_set_mesh_params gives the mesh other step parameters. */
int gsh_server_add_mesh(gsh_server* srv, gsh_mesh* mesh, int32_t* mesh_id_out);
/* Step parameters (dt, theta, iettype, ...) of a registered mesh, used by every request on it
 * admitted after the call; validated here (GSH_E_ARG). The reference runs one step parameter
 * set per model (schism_step.py), which a caller passes once per mesh. */
int gsh_server_set_mesh_params(gsh_server* srv, int32_t mesh_id, const gsh_step_params* p);
/* Submits one solve (+ optional backtrack) request: eta2/b host vectors are copied at submit.
 * with_backtrack != 0 runs gsh_backtrack on the mesh's current velocities before the solve.
 * b_host = NULL assembles the RHS from the mesh state. */
int gsh_server_submit(gsh_server* srv, int32_t mesh_id, const double* b_host, int with_backtrack,
                      int64_t* ticket_out);
/* Advances the service: admits queued requests as memory and slots allow (their device set-up,
 * i.e. backtracking, assembly and casts, runs on a second stream during this step's solve window
 * and they join at the next step), runs one check window of lock-step PCG iterations for every
 * active request, retires finished ones. active_out counts active and admitted requests. */
int gsh_server_step(gsh_server* srv, int32_t* active_out, int32_t* queued_out);
/* Tickets of the requests retired since the last call (up to cap, oldest first), so a caller
 * polls only finished requests. Requests rejected at admission are reported by gsh_server_poll. */
int gsh_server_finished(gsh_server* srv, int64_t* tickets, int32_t cap, int32_t* count);
/* Result of a finished request: status 0 = done, 1 = pending; x_host may be NULL. A finished
 * request keeps its solution in page-locked host memory until gsh_server_release. */
int gsh_server_poll(gsh_server* srv, int64_t ticket, double* x_host, int32_t* iters_out,
                    int32_t* status_out, double* latency_ms_out);
/* Forgets a finished or rejected request: its page-locked host vectors return to the server's
 * per-mesh pool for later requests, and the ticket becomes unknown (GSH_E_NOTFOUND).
 * GSH_E_STATE while the request is queued or running. No counterpart in the reference (it
 * has no server). */
int gsh_server_release(gsh_server* srv, int64_t ticket);

/* ---------------------------------------------------------------- measurement */
/* No counterpart in the reference: measurement helpers of this library.
 * gsh_copy_probe: average time of one device-to-device copy of `bytes` (read + write) by a
 * streaming kernel, the fastest of seven launch shapes (grid-stride double4 copies at 8-32
 * workgroups per CU, 1 or 4 records in flight per thread, cached or nontemporal stores): the
 * achievable-bandwidth reference. gsh_solver_profile: one FP64 solve (same
 * arithmetic as gsh_solver_solve) with HIP events around the launches of every iteration;
 * ms[4] receives the mean time per iteration of each launch kind:
 *   Jacobi: ms[0] phase A (p update + SpMV + p.q), ms[1] phase B (x, r, z updates + r.r, r.z);
 *   SSOR:   ms[0] backward sweeps (all colours, with the vector updates), ms[1] forward sweeps. */
int gsh_copy_probe(gsh_ctx* ctx, int64_t bytes, int reps, double* ms_avg);
int gsh_solver_profile(gsh_solver* s, double rtol, int maxiter, int check_every, double* ms, int* iters_out);

/* ---------------------------------------------------------------- synthetic inputs */
/* Jittered-grid Delaunay channel: nx*ny nodes spaced dx, interior nodes
 * jittered uniformly by +-jitter*dx (boundary nodes only along the boundary), cells split and
 * Lawson-flipped to Delaunay. The west edge (x = 0) is the open boundary. */
typedef struct gsh_gen gsh_gen;
int gsh_gen_channel(int32_t nx, int32_t ny, double dx, double jitter, uint64_t seed, gsh_gen** out);
int gsh_gen_sizes(const gsh_gen* g, int32_t* n, int32_t* ne);
int gsh_gen_get(const gsh_gen* g, double* xy, int32_t* elnode, int8_t* open_bnd);
int gsh_gen_destroy(gsh_gen* g);
/* Log-normal bathymetry from a Gaussian random field (Gaussian covariance, correlation length
 * corr_cells grid cells) on the nx*ny generator grid, mapped to [hmin, hmax]. */
int gsh_gen_bathymetry(int32_t nx, int32_t ny, double corr_cells, double hmin, double hmax,
                       uint64_t seed, double* depth_out);
/* Divergence-free velocity from a stream function of nmodes Fourier modes, scaled so that
 * max |u| over the domain is umax, times the vertical shear (1 + shear*sigma_k); w = 0.
 * out is [npts][nvrt][ncomp] with ncomp 2 (u,v) or 3 (u,v,w). */
int gsh_gen_velocity(int32_t npts, const double* xy, int32_t nvrt, const double* sigma,
                     double lx, double ly, double dx, int32_t nmodes, double umax, double shear,
                     uint64_t seed, int32_t ncomp, double* out);
/* Normal deviates N(mean, std^2) from the library's xoshiro256** / Box-Muller stream. */
int gsh_gen_normal(uint64_t seed, int64_t count, double mean, double std, double* out);

#ifdef __cplusplus
}
#endif
#endif /* GSH_H */
