// api.hip -- the extern "C" boundary of libgsh: contexts, meshes, step operations, the PCG
// solver driver (single, batched, mixed precision) and the continuous-batching server.
//
// Host-side counterpart of the reference's PCGSolverMultiGPU (pcg_solver.py:538-1151) and its
// compiled loop NativePCG (pcg_loop.cpp:152-465): the iteration loop issues the two fused PCG
// phases per iteration on one HIP stream and reads back one small state record per check
// window, asynchronously, so the GPU never idles waiting for the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "abi_util.h"
#include "context.h"
#include "device.h"
#include "ssor_layout.h"

namespace gsh {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }


// Device allocations owned by one object, released together.
struct DevPool {
  // Bump allocator over zeroed chunks that double in size: an object with dozens of buffers
  // costs a handful of hipMalloc / hipMemset calls instead of one pair per buffer.
  std::vector<void*> ptrs;
  int64_t bytes = 0;
  char* cur = nullptr;
  size_t left = 0, next_chunk = (size_t)1 << 20;
  template <class T>
  T* alloc(int64_t count, bool zero = true) {
    (void)zero;  // chunk memory is zeroed when the chunk is allocated
    if (count <= 0) count = 1;
    const size_t nb = ((size_t)count * sizeof(T) + 255) & ~(size_t)255;
    if (nb > left) {
      const size_t cb = std::max(nb, next_chunk);
      void* p = nullptr;
      HIPCHK(hipMalloc(&p, cb));
      HIPCHK(hipMemset(p, 0, cb));
      // the fill runs on the null stream, which the library's non-blocking streams do not wait
      // for: it must be complete before a kernel on those streams writes the new buffer
      HIPCHK(hipStreamSynchronize(nullptr));
      ptrs.push_back(p);
      cur = static_cast<char*>(p);
      left = cb;
      next_chunk = std::min<size_t>(2 * next_chunk, (size_t)1 << 30);
    }
    T* out = reinterpret_cast<T*>(cur);
    cur += nb;
    left -= nb;
    bytes += (int64_t)nb;
    return out;
  }
  template <class T>
  T* upload(const std::vector<T>& v) {
    T* p = alloc<T>((int64_t)v.size(), false);
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    bytes = 0;
    cur = nullptr;
    left = 0;
  }
  ~DevPool() { release(); }
};

template <class VT>
__global__ void k_gather_states(const PcgSys<VT>* __restrict__ sys, int nsys, PcgState* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsys) out[i] = *sys[i].st;
}

__global__ void k_extrapolate(double* __restrict__ x, const double* __restrict__ h0, const double* __restrict__ h1,
                              const double* __restrict__ h2, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (3.0 * h0[i] - 3.0 * h1[i]) + h2[i];
}

__global__ void k_pack_uvw(const double* __restrict__ in, double* __restrict__ out, int64_t cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  out[4 * i] = in[3 * i];
  out[4 * i + 1] = in[3 * i + 1];
  out[4 * i + 2] = in[3 * i + 2];
  out[4 * i + 3] = 0.0;
}

__global__ void k_unpack_uvw(const double* __restrict__ in, double* __restrict__ out, int64_t cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  out[3 * i] = in[4 * i];
  out[3 * i + 1] = in[4 * i + 1];
  out[3 * i + 2] = in[4 * i + 2];
}

}  // namespace gsh

using namespace gsh;

// ====================================================================== context
// Stream-ordered temporaries of one call, freed on the stream when the scope ends, on the error
// path as well.
struct StreamBuffers {
  hipStream_t st;
  std::vector<void*> ptrs;
  explicit StreamBuffers(hipStream_t s) : st(s) {}
  template <class T>
  T* alloc(size_t count) {
    void* p = nullptr;
    HIPCHK(hipMallocAsync(&p, sizeof(T) * std::max<size_t>(1, count), st));
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
  ~StreamBuffers() {
    for (void* p : ptrs) (void)hipFreeAsync(p, st);
  }
  StreamBuffers(const StreamBuffers&) = delete;
  StreamBuffers& operator=(const StreamBuffers&) = delete;
};

// ====================================================================== mesh
struct gsh_mesh {
  gsh_ctx* ctx = nullptr;
  HostMesh hm;
  Topology topo;
  Geometry geo;
  DMesh d{};
  DevPool pool;
  int32_t n_open = 0;
  // tidal forcing tables (gsh_set_tides): capacities of the pool buffers
  double *tide_glob = nullptr, *tide_emo = nullptr, *tide_efa = nullptr;
  int64_t tide_cap = 0;
  int32_t tide_nb = 0;
  double* shap_tmp = nullptr;  // Shapiro filter output before it replaces u2 (first use)
};

// ====================================================================== window graphs
// A window of PCG launches (check_every iterations and the window-end test) captured once per
// solver and replayed with one hipGraphLaunch: the kernels read the system descriptor from
// device memory, so one graph serves every solve of the solver. If capture is not available the
// window is launched directly.
struct WindowGraph {
  hipGraphExec_t exec = nullptr;
  const void* key = nullptr;
  int w = 0, rpt = 0;
  bool failed = false, enabled = false;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key = nullptr;
    w = 0;
  }
  ~WindowGraph() { reset(); }
};

// Window graphs are opt-in per solver (gsh_solver_set_graphs): on the 1M-node solves they
// measured no faster than direct launches, and a profiler's kernel tracing does not follow graph
// capture here.
template <class F>
static bool launch_window_graph(WindowGraph& g, hipStream_t st, const void* key, int w, int rpt, F&& issue) {
  if (g.failed || !g.enabled) return false;
  if (!(g.exec && g.key == key && g.w == w && g.rpt == rpt)) {
    g.reset();
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      g.failed = true;
      return false;
    }
    issue();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 =
        (e == hipSuccess && graph) ? hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) : hipErrorUnknown;
    if (graph) (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
      (void)hipGetLastError();
      g.failed = true;
      return false;
    }
    g.exec = exec;
    g.key = key;
    g.w = w;
    g.rpt = rpt;
  }
  HIPCHK(hipGraphLaunch(g.exec, st));
  return true;
}

// ====================================================================== solver
// Device side of the multicolour SSOR solver (built from the solver's CSR pattern).
struct SsorState {
  DevPool pool;
  SsorLayout L;
  SsorSys hs{};                 // FP64 solver
  SsorSys* ds = nullptr;
  SsorSysT<float> hs32{};       // mixed precision: the FP32 inner solver
  SsorSysT<float>* ds32 = nullptr;
  SsorSysT<float> up32{};       // the FP32 descriptor as last uploaded to ds32
  bool up32_valid = false;
  int rpt = 1, ntiles = 0;
  std::vector<int> tile0, ntc;  // tiles of each colour
};

struct gsh_solver {
  gsh_ctx* ctx = nullptr;
  gsh_mesh* mesh = nullptr;
  int precision = GSH_PREC_FP64;
  int rpt = 4;                 // rows per thread of a tile
  double inner_rtol = 1e-5;    // MIXED: relative tolerance of each inner FP32 solve
  int32_t n = 0, ntiles = 0;
  int64_t nnz = 0;
  DevPool pool;
  int32_t *rowptr = nullptr, *col = nullptr, *diag = nullptr;
  int16_t* cdel = nullptr;
  double *val = nullptr, *b = nullptr, *x = nullptr, *h[3] = {nullptr, nullptr, nullptr};
  PcgSys<double> hs{};         // host copy of the FP64 system descriptor
  PcgSys<double>* ds = nullptr;
  TileRef* dtiles = nullptr;   // tiles of this system alone
  double* red2 = nullptr;      // scratch for two reduced scalars
  // mixed precision
  float* val32 = nullptr;
  double* r64 = nullptr;
  PcgSys<float> hs32{};
  PcgSys<float>* ds32 = nullptr;
  gsh_solve_stats last{};
  // preconditioner
  int precond = GSH_PRECOND_JACOBI;
  double omega = 1.0;
  std::vector<int32_t> h_indptr, h_indices;  // host pattern kept for the SSOR layout
  std::unique_ptr<SsorState> ssor;
  // captured windows (declared after the memory they launch on, so they are destroyed first)
  WindowGraph wg_jacobi, wg_ssor;

  int tiles_for(int rows) const { return (rows + kTpb * rpt - 1) / (kTpb * rpt); }
};

static void solver_alloc(gsh_solver* s, int32_t n, int64_t nnz, const std::vector<int32_t>& indptr,
                         const std::vector<int32_t>& indices, const std::vector<int32_t>& diag_abs) {
  s->n = n;
  s->nnz = nnz;
  s->ntiles = s->tiles_for(n);
  auto& P = s->pool;
  const int nt = std::max(1, (n + kTpb - 1) / kTpb);  // room for the smallest tile size
  s->rowptr = P.upload(indptr);
  s->col = P.upload(indices);
  s->diag = P.upload(diag_abs);
  // int16 column offsets from the first row of each 64-row chunk, when the bandwidth allows
  {
    std::vector<int16_t> cdel(nnz);
    bool ok = true;
    for (int32_t i = 0; i < n && ok; ++i)
      for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
        const int32_t d = indices[k] - (i & ~63);
        if (d < -32768 || d > 32767) {
          ok = false;
          break;
        }
        cdel[k] = (int16_t)d;
      }
    s->cdel = ok ? P.upload(cdel) : nullptr;
  }
  s->val = P.alloc<double>(nnz);
  s->b = P.alloc<double>(n);
  s->x = P.alloc<double>(n);
  for (auto& hh : s->h) hh = P.alloc<double>(n);
  PcgSys<double>& h = s->hs;
  h.n = n;
  h.ntiles = s->ntiles;
  h.rowptr = s->rowptr;
  h.col = s->col;
  h.val = s->val;
  h.diag = s->diag;
  h.cdel = s->cdel;
  h.b = s->b;
  h.x = s->x;
  h.r = P.alloc<double>(n);
  h.z = P.alloc<double>(n);
  h.pa = P.alloc<double>(n);
  h.pb = P.alloc<double>(n);
  h.q = P.alloc<double>(n);
  h.invd = P.alloc<double>(n);
  h.part_rr = P.alloc<double>(nt);
  h.part_rz = P.alloc<double>(nt);
  h.part_pq = P.alloc<double>(nt);
  h.part_bb = P.alloc<double>(nt);
  h.st = P.alloc<PcgState>(1);
  s->ds = P.alloc<PcgSys<double>>(1);
  s->dtiles = P.alloc<TileRef>(nt);
  s->red2 = P.alloc<double>(2);
  if (s->precision == GSH_PREC_MIXED) {
    s->val32 = P.alloc<float>(nnz);
    s->r64 = P.alloc<double>(n);
    PcgSys<float>& g = s->hs32;
    g.n = n;
    g.ntiles = s->ntiles;
    g.rowptr = s->rowptr;
    g.col = s->col;
    g.val = s->val32;
    g.diag = s->diag;
    g.cdel = s->cdel;
    g.b = s->r64;
    g.x = P.alloc<float>(n);
    g.r = P.alloc<float>(n);
    g.z = P.alloc<float>(n);
    g.pa = P.alloc<float>(n);
    g.pb = P.alloc<float>(n);
    g.q = P.alloc<float>(n);
    g.invd = P.alloc<float>(n);
    g.h0 = g.h1 = g.h2 = nullptr;
    g.part_rr = P.alloc<double>(nt);
    g.part_rz = P.alloc<double>(nt);
    g.part_pq = P.alloc<double>(nt);
    g.part_bb = P.alloc<double>(nt);
    g.st = P.alloc<PcgState>(1);
    s->ds32 = P.alloc<PcgSys<float>>(1);
  }
}

static std::vector<TileRef> tile_table(int sys, int ntiles) {
  std::vector<TileRef> t(ntiles);
  for (int i = 0; i < ntiles; ++i) t[i] = {sys, i};
  return t;
}

// Re-uploads the descriptors and the single-system tile table (pointers rotate between solves).
static void solver_sync_desc(gsh_solver* s) {
  hipStream_t st = s->ctx->stream;
  s->hs.x = s->x;
  s->hs.h0 = s->h[0];
  s->hs.h1 = s->h[1];
  s->hs.h2 = s->h[2];
  s->hs.ntiles = s->ntiles;
  HIPCHK(hipMemcpyAsync(s->ds, &s->hs, sizeof(s->hs), hipMemcpyHostToDevice, st));
  if (s->ds32) {
    s->hs32.ntiles = s->ntiles;
    HIPCHK(hipMemcpyAsync(s->ds32, &s->hs32, sizeof(s->hs32), hipMemcpyHostToDevice, st));
  }
  auto t = tile_table(0, s->ntiles);
  HIPCHK(hipMemcpyAsync(s->dtiles, t.data(), t.size() * sizeof(TileRef), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // host staging buffers above are stack/heap temporaries
}

// Colour-major layout, split matrix and work vectors of the SSOR solver for the current
// rows-per-thread (the colour blocks are padded to whole tiles of 256 * rpt rows). An FP64
// solver gets FP64 values and vectors; a mixed-precision solver gets FP32 ones for its inner
// solves (the FP64 residual and refinement use the solver's CSR in the original order).
static void ssor_build(gsh_solver* s) {
  auto S = std::make_unique<SsorState>();
  S->rpt = s->rpt;
  const int T = kTpb * s->rpt;
  build_ssor_layout(s->n, s->h_indptr.data(), s->h_indices.data(), T, S->L);
  const SsorLayout& L = S->L;
  auto& P = S->pool;
  std::vector<SsorColour> cols(L.ncol);
  for (int c = 0; c < L.ncol; ++c)
    cols[c] = {L.cstart[c], L.cstart[c + 1] - L.cstart[c], L.wl[c], L.wu[c], (int32_t)L.offl[c], (int32_t)L.offu[c], 0, 0};
  auto fill = [&](auto& h) {
    using VT = std::remove_pointer_t<decltype(h.lval)>;
    h.n = s->n;
    h.np = L.np;
    h.ntiles = L.np / T;
    h.ncol = L.ncol;
    h.omega = s->omega;
    h.colours = P.upload(cols);
    h.tile_colour = P.upload(L.tile_colour);
    h.lcol = P.upload(L.lcol);
    h.lsrc = P.upload(L.lsrc);
    h.ucol = P.upload(L.ucol);
    h.usrc = P.upload(L.usrc);
    h.dsrc = P.upload(L.dsrc);
    h.perm = P.upload(L.perm);
    h.lval = P.alloc<VT>((int64_t)L.lcol.size());
    h.uval = P.alloc<VT>((int64_t)L.ucol.size());
    for (VT** v : {&h.d, &h.dt, &h.kd, &h.bp, &h.y, &h.r, &h.p, &h.q, &h.t, &h.u, &h.s, &h.w}) *v = P.alloc<VT>(L.np);
    for (double** v : {&h.part_rr, &h.part_bb}) *v = P.alloc<double>(h.ntiles);
    for (double** v : {&h.part_rz, &h.part_pq}) *v = P.alloc<double>(2 * (int64_t)h.ntiles);  // two iteration slots
    h.pre = P.alloc<double>(4 * (int64_t)kTpb);
    h.pre_t0 = L.cstart[L.ncol - 1] / T;
    h.st = P.alloc<PcgState>(1);
  };
  if (s->precision == GSH_PREC_MIXED) {
    fill(S->hs32);
    S->ds32 = P.alloc<SsorSysT<float>>(1);
    S->ntiles = S->hs32.ntiles;
  } else {
    fill(S->hs);
    S->ds = P.alloc<SsorSys>(1);
    S->ntiles = S->hs.ntiles;
  }
  for (int c = 0; c < L.ncol; ++c) {
    S->tile0.push_back(L.cstart[c] / T);
    S->ntc.push_back((L.cstart[c + 1] - L.cstart[c]) / T);
  }
  s->ssor = std::move(S);
}

template <class VT>
static SsorSysT<VT>& ssor_host(SsorState& S);
template <>
SsorSysT<double>& ssor_host<double>(SsorState& S) {
  return S.hs;
}
template <>
SsorSysT<float>& ssor_host<float>(SsorState& S) {
  return S.hs32;
}
template <class VT>
static SsorSysT<VT>* ssor_dev(SsorState& S);
template <>
SsorSysT<double>* ssor_dev<double>(SsorState& S) {
  return S.ds;
}
template <>
SsorSysT<float>* ssor_dev<float>(SsorState& S) {
  return S.ds32;
}

// FP64: the solver's values, b, x and history. FP32 (inner solve of mixed precision): the FP32
// copy of the values, the given FP64 right-hand side (the outer residual), the FP32 correction
// as solution and a zero initial guess.
// With a stream given the copy is only queued there (the host descriptor lives in the solver).
static void ssor_sync_desc(gsh_solver* s, hipStream_t on = nullptr) {
  SsorSys& h = s->ssor->hs;
  h.aval = s->val;
  h.b = s->b;
  h.xout = s->x;
  h.h0 = s->h[0];
  h.h1 = s->h[1];
  h.h2 = s->h[2];
  h.omega = s->omega;
  h.unit_omega = s->omega == 1.0;
  HIPCHK(hipMemcpyAsync(s->ssor->ds, &h, sizeof(h), hipMemcpyHostToDevice, on ? on : s->ctx->stream));
  if (!on) HIPCHK(hipStreamSynchronize(s->ctx->stream));
}

static void ssor_sync_desc32(gsh_solver* s, const double* rhs) {
  SsorSysT<float>& h = s->ssor->hs32;
  h.aval = s->val32;
  h.b = rhs;
  h.xout = s->hs32.x;
  h.h0 = h.h1 = h.h2 = nullptr;
  h.omega = s->omega;
  h.unit_omega = s->omega == 1.0;
  // the inner solves of one mixed solve share the descriptor: upload it only when it changed
  if (s->ssor->up32_valid && std::memcmp(&s->ssor->up32, &h, sizeof(h)) == 0) return;
  HIPCHK(hipMemcpyAsync(s->ssor->ds32, &h, sizeof(h), hipMemcpyHostToDevice, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  s->ssor->up32 = h;
  s->ssor->up32_valid = true;
}

// rollHistory (pcg_solver.py:941-947) by pointer rotation: x2 <- x1 <- x0 <- x.
static void solver_roll(gsh_solver* s) {
  double* old2 = s->h[2];
  s->h[2] = s->h[1];
  s->h[1] = s->h[0];
  s->h[0] = s->x;
  s->x = old2;
}

// ---------------------------------------------------------------------- PCG driver
struct PcgOutcome {
  std::vector<PcgState> st;
};

// When the host issues the windows of a solve. The host reads each window's state (the true
// r.r of its closing check) while the next windows run, so the windows issued behind the one
// that converges run kernels that return at once: up to two windows of empty launches per
// solve. From the last two checks read, the pacer extrapolates log(r.r) per window for every
// system still running; when all of them are expected to pass by the newest window in flight,
// the host waits for the windows in flight before it issues another (one GPU round trip
// instead of a window of empty launches). Pacing decides when windows are issued, never which
// iterations run or where the checks fall, so results do not depend on it.
struct WindowPacer {
  struct Sys {
    double l0 = 0.0, l1 = 0.0, la = 0.0;  // log r.r of the last two checks read, log atol^2
    int w0 = -1, w1 = -1;                 // their window indices
    bool fin = false;
  };
  Sys one;               // the state of a single system (no heap allocation)
  std::vector<Sys> many;  // several systems
  Sys* sys;
  int n;
  explicit WindowPacer(int nsys) : many(nsys > 1 ? nsys : 0), sys(nsys > 1 ? many.data() : &one), n(nsys) {}
  WindowPacer(const WindowPacer&) = delete;
  WindowPacer& operator=(const WindowPacer&) = delete;
  // the states of window `win`, one per system
  void read(const PcgState* s, int win) {
    for (int i = 0; i < n; ++i) {
      Sys& y = sys[i];
      if (s[i].done >= 0) {
        y.fin = true;
      } else if (s[i].rr_last > 0.0 && s[i].atol2 > 0.0 && win > y.w1) {
        y.l0 = y.l1;
        y.w0 = y.w1;
        y.l1 = std::log(s[i].rr_last);
        y.w1 = win;
        y.la = std::log(s[i].atol2);
      }
    }
  }
  // every running system is expected to pass its check by window `newest` (half a window of
  // margin: a miss costs a round trip, a late stop a window)
  bool expect_done(int newest) const {
    for (int i = 0; i < n; ++i) {
      const Sys& y = sys[i];
      if (y.fin) continue;
      if (y.w0 < 0) return false;
      const double rate = (y.l0 - y.l1) / (double)(y.w1 - y.w0);
      if (!(rate > 0.0) || y.l1 - rate * ((double)(newest - y.w1) + 0.5) > y.la) return false;
    }
    return true;
  }
};

// Runs the PCG iteration for nsys systems described by dsys (device) over the tile table
// dtiles. Each system stops at its own first check with r.r <= atol^2 (pcg_loop.cpp:405-412).
// Per-phase launch timing collected with HIP events on the solver stream.
struct PhaseTimes {
  double ms_a = 0.0, ms_b = 0.0;
  int64_t launches = 0;
};

// The iteration loop itself: dgather (device, 2 * nsys states) stages the window readbacks and
// the final states land in the context's pinned slots 0..nsys-1. It allocates nothing when the
// pinned slots already hold 2 * nsys states and no profile is asked for.
template <class VT>
static void run_pcg_core(gsh_ctx* ctx, const PcgSys<VT>* dsys, int nsys, const TileRef* dtiles, int ntiles, int rpt,
                         double rtol, int maxiter, int check_every, double timeout_s, PcgState* dgather,
                         const TileRef* dinit_tiles, int ninit_tiles, PhaseTimes* prof, const double* rtols_dev,
                         WindowGraph* graph) {
  hipStream_t st = ctx->stream;
  if (check_every < 1) check_every = 1;
  if (timeout_s <= 0.0) timeout_s = 60.0;
  PcgState* slots = ctx->pinned_slots(nsys);
  launch_pcg_reset<VT>(dsys, nsys, rtol, maxiter, check_every, st, rtols_dev);
  if (dinit_tiles)
    launch_pcg_init<VT>(dsys, dinit_tiles, ninit_tiles, rpt, st);
  else
    launch_pcg_init<VT>(dsys, dtiles, ntiles, rpt, st);
  HIPCHK(hipGetLastError());
  int issued = 0, window = 0;
  int inflight[2], nflight = 0;  // windows pending readback (slot = window & 1), oldest first
  bool all_done = false;
  WindowPacer pacer(nsys);
  const auto t0 = Clock::now();
  auto check_oldest = [&]() {
    const int win = inflight[0], slot = win & 1;
    inflight[0] = inflight[1];
    --nflight;
    wait_event(ctx->ev[slot], timeout_s, "PCG window");
    bool done = true;
    for (int i = 0; i < nsys; ++i) done = done && (slots[slot * nsys + i].done >= 0);
    pacer.read(slots + slot * nsys, win);
    return done;
  };
  while (issued < maxiter && !all_done) {
    while (!all_done && nflight > 0 && pacer.expect_done(window - 1)) all_done = check_oldest();
    if (all_done) break;
    const int w = std::min(check_every, maxiter - issued);
    if (prof) {
      // timed window: the runtime stamps each phase kernel's start and end into its own event
      // pair; read back before the next window
      std::vector<hipEvent_t> evs(4 * w);
      for (auto& e : evs) HIPCHK(hipEventCreate(&e));
      for (int i = 0; i < w; ++i) {
        launch_pcg_phase<VT>(0, dsys, dtiles, ntiles, rpt, st, evs[4 * i], evs[4 * i + 1]);
        launch_pcg_phase<VT>(1, dsys, dtiles, ntiles, rpt, st, evs[4 * i + 2], evs[4 * i + 3]);
      }
      HIPCHK(hipEventSynchronize(evs[4 * w - 1]));
      for (int i = 0; i < w; ++i) {
        float a = 0.f, b = 0.f;
        HIPCHK(hipEventElapsedTime(&a, evs[4 * i], evs[4 * i + 1]));
        HIPCHK(hipEventElapsedTime(&b, evs[4 * i + 2], evs[4 * i + 3]));
        prof->ms_a += a;
        prof->ms_b += b;
        prof->launches += 1;
      }
      for (auto& e : evs) (void)hipEventDestroy(e);
      launch_pcg_check<VT>(dsys, nsys, st);
    } else {
      auto issue = [&]() {
        for (int i = 0; i < w; ++i) launch_pcg_iter<VT>(dsys, dtiles, ntiles, rpt, st);
        launch_pcg_check<VT>(dsys, nsys, st);
      };
      if (!(graph && launch_window_graph(*graph, st, dsys, w, rpt, issue))) issue();
    }
    issued += w;
    const int slot = window & 1;
    if (nflight == 2 && check_oldest()) all_done = true;
    hipLaunchKernelGGL(k_gather_states<VT>, dim3((nsys + 63) / 64), dim3(64), 0, st, dsys, nsys,
                       dgather + slot * nsys);
    HIPCHK(hipMemcpyAsync(slots + slot * nsys, dgather + slot * nsys, sizeof(PcgState) * nsys,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ctx->ev[slot], st));
    inflight[nflight++] = window;
    ++window;
    if (std::chrono::duration<double>(Clock::now() - t0).count() > timeout_s)
      throw Error(GSH_E_TIMEOUT, "PCG: time limit exceeded while issuing iterations");
    HIPCHK(hipGetLastError());
  }
  launch_pcg_finish<VT>(dsys, dtiles, ntiles, rpt, st);
  hipLaunchKernelGGL(k_gather_states<VT>, dim3((nsys + 63) / 64), dim3(64), 0, st, dsys, nsys, dgather);
  HIPCHK(hipMemcpyAsync(slots, dgather, sizeof(PcgState) * nsys, hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(ctx->ev[0], st));
  wait_event(ctx->ev[0], timeout_s, "PCG finish");
}

template <class VT>
static void run_pcg(gsh_ctx* ctx, const PcgSys<VT>* dsys, int nsys, const TileRef* dtiles, int ntiles, int rpt,
                    double rtol, int maxiter, int check_every, double timeout_s, PcgOutcome& out,
                    const TileRef* dinit_tiles = nullptr, int ninit_tiles = -1, PhaseTimes* prof = nullptr,
                    const double* rtols_dev = nullptr, WindowGraph* graph = nullptr) {
  StreamBuffers tmp(ctx->stream);
  PcgState* dgather = tmp.alloc<PcgState>(2 * (size_t)nsys);
  run_pcg_core<VT>(ctx, dsys, nsys, dtiles, ntiles, rpt, rtol, maxiter, check_every, timeout_s, dgather, dinit_tiles,
                   ninit_tiles, prof, rtols_dev, graph);
  const PcgState* slots = ctx->pinned;
  out.st.assign(slots, slots + nsys);
}

static int iters_of(const PcgState& s, int maxiter) { return s.done >= 0 ? s.done : maxiter; }

// Per-iteration kernel time of the SSOR solver by launch kind (HIP events, profiling only).
struct SsorTimes {
  double ms_bwd = 0.0, ms_fwd = 0.0, ms_upd = 0.0;
  int64_t iters = 0;
};

// Flags of an SSOR sweep launch (the kernels' kFirst .. kLead): first sweep of an iteration,
// colour 0 (forms u and s), last sweep of an iteration, set-up pass, workgroup that writes the
// system's scalars (batched launches).
constexpr int kSsorFirst = 1, kSsorMid = 2, kSsorLast = 4, kSsorInit = 8, kSsorLead = 16, kSsorPre = 32,
              kSsorPreWG = 64, kSsorChk = 128;

// Launch table of a batch of SSOR systems: backward colour steps 0 .. maxC-1 at boff, forward
// steps 0 .. maxC-2 at foff (with each system's pre-sum entry in the step of its last forward
// sweep).
struct SsorTable {
  std::vector<SsorRef> refs;
  std::vector<int> boff, foff;
  int maxC = 0;
};

static SsorTable ssor_table(gsh_solver* const* S, int count) {
  SsorTable tb;
  for (int i = 0; i < count; ++i) tb.maxC = std::max(tb.maxC, S[i]->ssor->L.ncol);
  const int maxC = tb.maxC;
  auto& refs = tb.refs;
  tb.boff.assign(maxC + 1, 0);
  tb.foff.assign(std::max(1, maxC), 0);
  auto add_colour = [&](int i, int c, int flags) {
    const SsorState& T = *S[i]->ssor;
    for (int t = 0; t < T.ntc[c]; ++t) refs.push_back({i, T.tile0[c] + t, c, flags | (t == 0 ? kSsorLead : 0)});
  };
  // a system of two colours or more starts its first sweep from the sums that an extra entry
  // (tile -1, kSsorPreWG) forms in the step of its last forward sweep
  for (int step = 0; step < maxC; ++step) {
    tb.boff[step] = (int)refs.size();
    for (int i = 0; i < count; ++i) {
      const int C = S[i]->ssor->L.ncol;
      if (step >= C) continue;
      const int c = C - 1 - step;
      add_colour(i, c, (step == 0 ? kSsorFirst | (C >= 2 ? kSsorPre : 0) : 0) | (c == 0 ? kSsorMid : 0) |
                           (C == 1 ? kSsorLast : 0));
    }
  }
  tb.boff[maxC] = (int)refs.size();
  for (int step = 0; step + 1 < maxC; ++step) {
    tb.foff[step] = (int)refs.size();
    for (int i = 0; i < count; ++i) {
      const int C = S[i]->ssor->L.ncol, c = step + 1;
      if (c < C) add_colour(i, c, c == C - 1 ? kSsorLast : 0);
      if (c == C - 1) refs.push_back({i, -1, c, kSsorPreWG});
    }
  }
  tb.foff[std::max(0, maxC - 1)] = (int)refs.size();
  for (const SsorRef& r : refs)  // every entry names a tile of its own system, or is a pre-sum entry
    GSH_REQUIRE(r.sys >= 0 && r.sys < count &&
                    ((r.tile >= 0 && r.tile < S[r.sys]->ssor->ntiles) || (r.tile == -1 && r.flags == kSsorPreWG)),
                GSH_E_STATE, "bad batched SSOR table");
  return tb;
}

// One SSOR iteration (or, with kSsorInit, the set-up pass) of every system of a table.
template <class VT>
static void ssor_table_sweeps(const SsorSysT<VT>* dsys, const SsorRef* drefs, const SsorTable& tb, int or_flags,
                              int rpt, hipStream_t st) {
  for (int step = 0; step < tb.maxC; ++step)
    launch_ssor_bwd_batch(dsys, drefs + tb.boff[step], tb.boff[step + 1] - tb.boff[step], or_flags, rpt, st);
  for (int step = 0; step + 1 < tb.maxC; ++step)
    launch_ssor_fwd_batch(dsys, drefs + tb.foff[step], tb.foff[step + 1] - tb.foff[step], or_flags, rpt, st);
}

// Per-system start of an SSOR solve: state, ELL values (unless kept), r0 = (D~ + L)^-1 (b - A x0)
// and, with init_pass, the set-up sweeps (z0, s0, partials of iteration 0) by the one-system
// kernels, whose bodies are the batched ones.
template <class VT>
static void ssor_begin(gsh_solver* s, double rtol, int maxiter, int check_every, bool gather, bool init_pass,
                       hipStream_t st) {
  SsorState& T = *s->ssor;
  const SsorSysT<VT>* ds = ssor_dev<VT>(T);
  const int rpt = T.rpt, C = T.L.ncol;
  launch_ssor_reset(ds, rtol, maxiter, check_every, st);
  launch_ssor_setup(ds, T.ntiles, rpt, st, gather);
  launch_ssor_residual(ds, T.ntiles, rpt, st);
  for (int c = 0; c < C; ++c) launch_ssor_r0(ds, T.tile0[c], T.ntc[c], c, rpt, st);
  if (!init_pass) return;
  for (int c = C - 1; c >= 0; --c)
    launch_ssor_bwd(ds, T.tile0[c], T.ntc[c], c,
                    (c == C - 1 ? kSsorFirst : 0) | (c == 0 ? kSsorMid : 0) | (C == 1 ? kSsorLast : 0) | kSsorInit,
                    rpt, st);
  // the last forward sweep also forms the sums the first sweep of iteration 0 starts from
  for (int c = 1; c < C; ++c)
    launch_ssor_fwd(ds, T.tile0[c], T.ntc[c], c, (c == C - 1 ? kSsorLast | kSsorPreWG : 0) | kSsorInit, rpt, st);
}

// End of an SSOR solve: x = (D~ + U)^-1 y, back to the original row order.
template <class VT>
static void ssor_finish(gsh_solver* s, hipStream_t st) {
  SsorState& T = *s->ssor;
  const SsorSysT<VT>* ds = ssor_dev<VT>(T);
  for (int c = T.L.ncol - 1; c >= 0; --c) launch_ssor_xsolve(ds, T.tile0[c], T.ntc[c], c, T.rpt, st);
  launch_ssor_scatter(ds, T.ntiles, T.rpt, st);
}

// Multicolour SSOR PCG of one system: set-up, windows of check_every iterations (2 * ncolours
// launches each) closed by the true-residual test, then x = (D~ + U)^-1 y. The descriptor of
// the VT system must be synchronised first (ssor_sync_desc / ssor_sync_desc32).
template <class VT>
static PcgState run_ssor(gsh_solver* s, double rtol, int maxiter, int check_every, double timeout_s,
                         SsorTimes* prof = nullptr, bool use_graph = false, bool gather = true) {
  gsh_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  if (check_every < 1) check_every = 1;
  if (timeout_s <= 0.0) timeout_s = 60.0;
  SsorState& S = *s->ssor;
  const SsorSysT<VT>* ds = ssor_dev<VT>(S);
  const int rpt = S.rpt, C = S.L.ncol;
  PcgState* slots = ctx->pinned_slots(1);
  // set-up, then the set-up pass: z0 = D~ r0, s0 = A^ z0 and the gamma, delta partials of iteration 0
  ssor_begin<VT>(s, rtol, maxiter, check_every, gather, true, st);
  // flags of the sweeps (see ssor_kernels.hip): first sweep, colour 0, last of the iteration;
  // with two colours or more the last sweep (a forward one) forms the per-thread sums of the
  // gamma and delta partials that the next first sweep starts from
  auto bflags = [&](int c) {
    return (c == C - 1 ? kSsorFirst | (C >= 2 ? kSsorPre : 0) : 0) | (c == 0 ? kSsorMid : 0) |
           (C == 1 ? kSsorLast : 0);
  };
  auto fflags = [&](int c) { return c == C - 1 ? kSsorLast | kSsorPreWG : 0; };
  // chk: the iteration closes a window, so its sweeps also write the true-residual partials
  auto bwd = [&](int chk) {
    for (int c = C - 1; c >= 0; --c) launch_ssor_bwd(ds, S.tile0[c], S.ntc[c], c, bflags(c) | chk, rpt, st);
  };
  auto fwd = [&](int chk) {
    for (int c = 1; c < C; ++c) launch_ssor_fwd(ds, S.tile0[c], S.ntc[c], c, fflags(c) | chk, rpt, st);
  };
  HIPCHK(hipGetLastError());
  int issued = 0, window = 0;
  int inflight[2], nflight = 0;  // windows pending readback (slot = window & 1), oldest first
  bool done = false;
  WindowPacer pacer(1);
  const auto t0 = Clock::now();
  auto check_oldest = [&]() {
    const int win = inflight[0], slot = win & 1;
    inflight[0] = inflight[1];
    --nflight;
    wait_event(ctx->ev[slot], timeout_s, "SSOR PCG window");
    pacer.read(slots + slot, win);
    return slots[slot].done >= 0;
  };
  while (issued < maxiter && !done) {
    while (!done && nflight > 0 && pacer.expect_done(window - 1)) done = check_oldest();
    if (done) break;
    const int w = std::min(check_every, maxiter - issued);
    if (prof) {
      // every launch stamps its own kernel start and end (hipExtLaunchKernelGGL events)
      const int nl = 2 * C - 1;
      std::vector<hipEvent_t> e(2 * nl);
      for (auto& x : e) HIPCHK(hipEventCreate(&x));
      for (int i = 0; i < w; ++i) {
        int l = 0;
        const int chk = i == w - 1 ? kSsorChk : 0;
        for (int c = C - 1; c >= 0; --c, ++l)
          launch_ssor_bwd(ds, S.tile0[c], S.ntc[c], c, bflags(c) | chk, rpt, st, e[2 * l], e[2 * l + 1]);
        for (int c = 1; c < C; ++c, ++l)
          launch_ssor_fwd(ds, S.tile0[c], S.ntc[c], c, fflags(c) | chk, rpt, st, e[2 * l], e[2 * l + 1]);
        HIPCHK(hipEventSynchronize(e[2 * (l - 1) + 1]));
        for (int m = 0; m < l; ++m) {
          float a = 0.f;
          HIPCHK(hipEventElapsedTime(&a, e[2 * m], e[2 * m + 1]));
          (m < C ? prof->ms_bwd : prof->ms_fwd) += a;
        }
        prof->iters += 1;
      }
      for (auto& x : e) (void)hipEventDestroy(x);
      launch_ssor_check(ds, st);
    } else {
      auto issue = [&]() {
        for (int i = 0; i < w; ++i) {
          const int chk = i == w - 1 ? kSsorChk : 0;
          bwd(chk);
          fwd(chk);
        }
        launch_ssor_check(ds, st);
      };
      if (!(use_graph && launch_window_graph(s->wg_ssor, st, ds, w, rpt, issue))) issue();
    }
    issued += w;
    const int slot = window & 1;
    if (nflight == 2 && check_oldest()) done = true;
    HIPCHK(hipMemcpyAsync(slots + slot, ssor_host<VT>(S).st, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ctx->ev[slot], st));
    inflight[nflight++] = window;
    ++window;
    if (std::chrono::duration<double>(Clock::now() - t0).count() > timeout_s)
      throw Error(GSH_E_TIMEOUT, "SSOR PCG: time limit exceeded while issuing iterations");
    HIPCHK(hipGetLastError());
  }
  ssor_finish<VT>(s, st);
  HIPCHK(hipMemcpyAsync(slots, ssor_host<VT>(S).st, sizeof(PcgState), hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(ctx->ev[0], st));
  wait_event(ctx->ev[0], timeout_s, "SSOR PCG finish");
  HIPCHK(hipGetLastError());
  return slots[0];
}

// Multicolour SSOR PCG of `count` systems in lock-step: the per-system set-up and the final
// x = (D~ + U)^-1 y as in run_ssor, the sweeps of all systems batched into one launch per colour
// step over a table of (system, tile, colour, flags) entries (a system with fewer colours sits
// out the steps beyond its own). Every system stops at its own first passing test; its arithmetic
// does not depend on the others, so each result equals its solo solve. The VT descriptors of the
// solvers must be synchronised first; rtols are per system; gather[i] = 0 keeps system i's ELL
// values from its previous set-up.
template <class VT>
static void run_ssor_batch(gsh_solver* const* S, int count, const std::vector<double>& rtols, int maxiter,
                           int check_every, double timeout_s, const std::vector<char>& gather,
                           std::vector<PcgState>& out) {
  gsh_ctx* ctx = S[0]->ctx;
  hipStream_t st = ctx->stream;
  if (check_every < 1) check_every = 1;
  if (timeout_s <= 0.0) timeout_s = 60.0;
  const int rpt = S[0]->ssor->rpt;
  std::vector<SsorSysT<VT>> hsys(count);
  for (int i = 0; i < count; ++i) {
    SsorState& T = *S[i]->ssor;
    GSH_REQUIRE(T.rpt == rpt, GSH_E_ARG, "batched SSOR solvers need the same rows per thread");
    hsys[i] = ssor_host<VT>(T);
  }
  const SsorTable tb = ssor_table(S, count);
  StreamBuffers tmp(st);
  SsorSysT<VT>* dsys = tmp.alloc<SsorSysT<VT>>(count);
  SsorRef* drefs = tmp.alloc<SsorRef>(tb.refs.size());
  PcgState* dgather = tmp.alloc<PcgState>(2 * (size_t)count);
  HIPCHK(hipMemcpyAsync(dsys, hsys.data(), sizeof(SsorSysT<VT>) * count, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(drefs, tb.refs.data(), sizeof(SsorRef) * tb.refs.size(), hipMemcpyHostToDevice, st));
  for (int i = 0; i < count; ++i) ssor_begin<VT>(S[i], rtols[i], maxiter, check_every, gather[i] != 0, false, st);
  ssor_table_sweeps(dsys, drefs, tb, kSsorInit, rpt, st);  // z0, s0 and the partials of iteration 0
  HIPCHK(hipGetLastError());
  PcgState* slots = ctx->pinned_slots(count);
  int issued = 0, window = 0;
  int inflight[2], nflight = 0;  // windows pending readback (slot = window & 1), oldest first
  bool all_done = false;
  WindowPacer pacer(count);
  const auto t0 = Clock::now();
  auto check_oldest = [&]() {
    const int win = inflight[0], slot = win & 1;
    inflight[0] = inflight[1];
    --nflight;
    wait_event(ctx->ev[slot], timeout_s, "batched SSOR PCG window");
    bool done = true;
    for (int i = 0; i < count; ++i) done = done && (slots[slot * count + i].done >= 0);
    pacer.read(slots + slot * count, win);
    return done;
  };
  while (issued < maxiter && !all_done) {
    while (!all_done && nflight > 0 && pacer.expect_done(window - 1)) all_done = check_oldest();
    if (all_done) break;
    const int w = std::min(check_every, maxiter - issued);
    // the window's last iteration also writes the true-residual partials
    for (int it = 0; it < w; ++it) ssor_table_sweeps(dsys, drefs, tb, it == w - 1 ? kSsorChk : 0, rpt, st);
    launch_ssor_check_batch(dsys, count, st);
    issued += w;
    const int slot = window & 1;
    if (nflight == 2 && check_oldest()) all_done = true;
    launch_ssor_gather_states(dsys, count, dgather + slot * count, st);
    HIPCHK(hipMemcpyAsync(slots + slot * count, dgather + slot * count, sizeof(PcgState) * count,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ctx->ev[slot], st));
    inflight[nflight++] = window;
    ++window;
    if (std::chrono::duration<double>(Clock::now() - t0).count() > timeout_s)
      throw Error(GSH_E_TIMEOUT, "batched SSOR PCG: time limit exceeded while issuing iterations");
    HIPCHK(hipGetLastError());
  }
  for (int i = 0; i < count; ++i) ssor_finish<VT>(S[i], st);
  launch_ssor_gather_states(dsys, count, dgather, st);
  HIPCHK(hipMemcpyAsync(slots, dgather, sizeof(PcgState) * count, hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(ctx->ev[0], st));
  wait_event(ctx->ev[0], timeout_s, "batched SSOR PCG finish");
  HIPCHK(hipGetLastError());
  out.assign(slots, slots + count);
}

// Mixed precision for `count` systems in lock-step: FP64 true residual and refinement around FP32
// PCG inner solves (FP64 dot products). Every system keeps its own outer state: its inner
// tolerance is max(inner_rtol, 0.5 sqrt(atol^2 / r.r)) from its own residual, each inner solve is
// bounded by maxiter, and the inner iterations add up to at most maxiter plus one inner solve,
// so a system's result does not depend on the other systems of the batch.
static void run_mixed(gsh_ctx* ctx, gsh_solver* const* S, int count, double rtol, int maxiter, int check_every,
                      double timeout_s, std::vector<int>& iters, std::vector<int>& status) {
  hipStream_t st = ctx->stream;
  const int rpt = S[0]->rpt;
  std::vector<PcgSys<double>> h64(count);
  std::vector<PcgSys<float>> h32(count);
  for (int i = 0; i < count; ++i) {
    gsh_solver* s = S[i];
    launch_cast_values(s->val, s->val32, s->nnz, st);
    hipLaunchKernelGGL(k_extrapolate, dim3((s->n + 255) / 256), dim3(256), 0, st, s->x, s->h[0], s->h[1], s->h[2],
                       s->n);
    h64[i] = s->hs;
    h64[i].x = s->x;
    h64[i].r = s->r64;
    h64[i].ntiles = s->ntiles;
    h32[i] = s->hs32;
    h32[i].ntiles = s->ntiles;
  }
  int64_t all_tiles = 0;
  for (int i = 0; i < count; ++i) all_tiles += S[i]->ntiles;
  StreamBuffers tmp(st);
  PcgSys<double>* d64 = tmp.alloc<PcgSys<double>>(count);
  PcgSys<double>* d64s = tmp.alloc<PcgSys<double>>(count);
  PcgSys<float>* d32s = tmp.alloc<PcgSys<float>>(count);
  TileRef* dtl = tmp.alloc<TileRef>((size_t)all_tiles);
  double* dred = tmp.alloc<double>(2 * (size_t)count);
  double* drt = tmp.alloc<double>(count);
  HIPCHK(hipMemcpyAsync(d64, h64.data(), sizeof(PcgSys<double>) * count, hipMemcpyHostToDevice, st));
  std::vector<double> atol2(count, -1.0), rr(count, 0.0), red(2 * count);
  std::vector<int> outer(count, 0), inner(count, 0), state(count, 0);  // 0 running, 1 converged, 2 stopped
  const int max_outer = 30;
  std::vector<char> ssor_gathered(count, 0);  // SSOR inner solves: gathered by each system's first
  const bool ssor = S[0]->precond == GSH_PRECOND_SSOR;
  TileRef* dres = ssor ? tmp.alloc<TileRef>((size_t)all_tiles) : dtl;  // residual tiles (SSOR: kept)
  std::vector<TileRef> tl, tl_up;  // residual tiles of the running systems; as last uploaded to dres
  for (int round = 0; round < max_outer; ++round) {
    // FP64 true residual of every running system
    tl.clear();
    for (int i = 0; i < count; ++i)
      if (state[i] == 0)
        for (int t = 0; t < S[i]->ntiles; ++t) tl.push_back({i, t});
    if (tl.empty()) break;
    if (!ssor || tl.size() != tl_up.size() ||
        std::memcmp(tl.data(), tl_up.data(), sizeof(TileRef) * tl.size()) != 0) {
      HIPCHK(hipMemcpyAsync(dres, tl.data(), sizeof(TileRef) * tl.size(), hipMemcpyHostToDevice, st));
      tl_up = tl;
    }
    launch_residual_batch(d64, dres, (int)tl.size(), rpt, st);
    launch_reduce_systems(d64, count, dred, st);
    HIPCHK(hipMemcpyAsync(red.data(), dred, sizeof(double) * 2 * count, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    wait_event(ctx->ev[0], timeout_s, "refinement residual");
    // outer decisions; the running systems that need an inner solve, in index order
    std::vector<int> sub;
    std::vector<double> rts;
    for (int i = 0; i < count; ++i) {
      if (state[i] != 0) continue;
      rr[i] = red[2 * i];
      outer[i] += 1;
      if (round == 0) {
        const double atol = rtol * std::sqrt(red[2 * i + 1]);
        atol2[i] = atol * atol;
        if (atol == 0.0) {  // pcg_solver.py:1089-1095
          launch_fill_f64(S[i]->x, 0.0, S[i]->n, st);
          state[i] = 1;
          continue;
        }
      }
      if (rr[i] <= atol2[i]) {
        state[i] = 1;
      } else if (inner[i] >= maxiter) {
        state[i] = 2;
      } else {
        sub.push_back(i);
        rts.push_back(std::max(S[i]->inner_rtol, 0.5 * std::sqrt(atol2[i] / rr[i])));
      }
    }
    if (sub.empty()) break;
    // inner FP32 solves A d = r of the subset, lock-step, from d = 0
    const int ns = (int)sub.size();
    PcgOutcome o;
    if (ssor) {
      // FP32 SSOR PCG of A d = r for the subset, b = each system's FP64 residual; the first inner
      // solve of a system gathers its FP32 matrix, the later ones reuse it
      std::vector<gsh_solver*> ss(ns);
      std::vector<char> gath(ns);
      for (int j = 0; j < ns; ++j) {
        ss[j] = S[sub[j]];
        ssor_sync_desc32(ss[j], ss[j]->r64);
        gath[j] = ssor_gathered[sub[j]] ? 0 : 1;
        ssor_gathered[sub[j]] = 1;
      }
      if (ns == 1)  // the one-system kernels (same bodies, no launch table)
        o.st.assign(1, run_ssor<float>(ss[0], rts[0], maxiter, check_every, timeout_s, nullptr, false, gath[0] != 0));
      else
        run_ssor_batch<float>(ss.data(), ns, rts, maxiter, check_every, timeout_s, gath, o.st);
      // x += d, d the FP32 correction in the original row order
      for (int j = 0; j < ns; ++j) launch_refine_update(ss[j]->x, ss[j]->hs32.x, ss[j]->n, st);
    } else {
      std::vector<PcgSys<double>> s64(ns);
      std::vector<PcgSys<float>> s32(ns);
      std::vector<TileRef> stl;
      for (int j = 0; j < ns; ++j) {
        s64[j] = h64[sub[j]];
        s32[j] = h32[sub[j]];
        for (int t = 0; t < S[sub[j]]->ntiles; ++t) stl.push_back({j, t});
      }
      HIPCHK(hipMemcpyAsync(d64s, s64.data(), sizeof(PcgSys<double>) * ns, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(d32s, s32.data(), sizeof(PcgSys<float>) * ns, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(dtl, stl.data(), sizeof(TileRef) * stl.size(), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(drt, rts.data(), sizeof(double) * ns, hipMemcpyHostToDevice, st));
      run_pcg<float>(ctx, d32s, ns, dtl, (int)stl.size(), rpt, rtol, maxiter, check_every, timeout_s, o, nullptr, -1,
                     nullptr, drt);
      launch_refine_batch(d64s, d32s, dtl, (int)stl.size(), rpt, st);
    }
    for (int j = 0; j < ns; ++j) {
      const int it = iters_of(o.st[j], maxiter);
      inner[sub[j]] += it;
      if (it == 0) state[sub[j]] = 2;  // no progress possible
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  iters.assign(count, 0);
  status.assign(count, 0);
  for (int i = 0; i < count; ++i) {
    gsh_solver* s = S[i];
    iters[i] = inner[i];
    status[i] = state[i] == 1 ? 0 : GSH_W_MAXITER;
    s->last.iters = inner[i];
    s->last.outer = outer[i];
    s->last.rr = rr[i];
    s->last.atol2 = atol2[i];
  }
}

static void solve_mixed(gsh_solver* s, double rtol, int maxiter, int check_every, double timeout_s, int* iters_out,
                        int* status_out) {
  std::vector<int> it, stv;
  gsh_solver* one[1] = {s};
  run_mixed(s->ctx, one, 1, rtol, maxiter, check_every, timeout_s, it, stv);
  *iters_out = it[0];
  *status_out = stv[0];
}

// ====================================================================== C ABI
// ====================================================================== solve without allocation
// gsh_pcg_solve_static keeps every buffer of its solve beyond the caller's inputs and output in
// one file-scope device array, `scratch`, laid out as below for up to GSH_STATIC_MAX_ROWS rows
// (about 53 MB, held by every device the library runs on until the process ends). Host code
// never passes `scratch` itself to the device: gsh_ctx_create looks up its address on each
// device once (hipGetSymbolAddress) and the entry point uses that address.
namespace gsh {
constexpr int kStaticRpt = 4;  // rows per thread of a fresh Jacobi solver (gsh_solver.rpt)
constexpr int64_t kStaticTiles = ((int64_t)GSH_STATIC_MAX_ROWS + kTpb - 1) / kTpb;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kScrSys = 0;                                                      // PcgSys<double>
constexpr size_t kScrState = kScrSys + align256(sizeof(PcgSys<double>));            // PcgState
constexpr size_t kScrGather = kScrState + align256(sizeof(PcgState));               // 2 PcgState
constexpr size_t kScrFlag = kScrGather + align256(2 * sizeof(PcgState));            // int32 check flag
constexpr size_t kScrTiles = kScrFlag + 256;                                        // TileRef per tile
constexpr size_t kScrPart = kScrTiles + align256(kStaticTiles * sizeof(TileRef));   // 4 x double per tile
constexpr size_t kScrDiag = kScrPart + 4 * align256(kStaticTiles * sizeof(double));  // int32 per row
constexpr size_t kScrVec = kScrDiag + align256((size_t)GSH_STATIC_MAX_ROWS * 4);     // r, z, pa, pb, q, invd
constexpr size_t kScrVecBytes = align256((size_t)GSH_STATIC_MAX_ROWS * sizeof(double));
constexpr size_t kScratchBytes = kScrVec + 6 * kScrVecBytes;
constexpr int kStaticMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static char* g_scratch_dev[kStaticMaxDevices] = {};  // address of `scratch` on each device

// With `device` current: finds the device address of `scratch` once per device.
static void static_scratch_setup(int device) {
  if (device < 0 || device >= kStaticMaxDevices) return;  // gsh_pcg_solve_static then refuses
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(GSH_E_DEVICE, "no device address for the static scratch array");
  g_scratch_dev[device] = static_cast<char*>(p);
}

__global__ void k_static_clear(int32_t* __restrict__ flag) { *flag = 0; }

// The caller's CSR, read on the device: flag bits 1 (indptr is not a non-decreasing sequence from
// 0 to nnz), 2 (a column outside 0..n-1), 4 (a row without a diagonal entry); diag[i] = index of
// row i's diagonal entry (the last one if repeated, as the solver's own set-up). Also writes the
// single-system tile table.
__global__ void k_static_csr(const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices, int32_t n,
                             int64_t nnz, int32_t* __restrict__ diag, int32_t* __restrict__ flag,
                             TileRef* __restrict__ tiles, int32_t ntiles) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntiles) tiles[i] = {0, i};
  if (i >= n) return;
  const int64_t a = indptr[i], b = indptr[i + 1];
  int f = 0;
  if ((i == 0 && a != 0) || (i == n - 1 && b != nnz)) f |= 1;
  int32_t d = -1;
  if (a < 0 || b < a || b > nnz) {
    f |= 1;
  } else {
    for (int64_t k = a; k < b; ++k) {
      const int32_t c = indices[k];
      if (c < 0 || c >= n)
        f |= 2;
      else if (c == i)
        d = (int32_t)k;
    }
    if (d < 0) f |= 4;
  }
  diag[i] = d;
  if (f) atomicOr(flag, f);
}

__global__ void k_static_desc(PcgSys<double> v, PcgSys<double>* __restrict__ out) { *out = v; }
}  // namespace gsh

extern "C" {

const char* gsh_last_error(void) { return g_last_error.c_str(); }
int gsh_version(void) { return 100; }

int gsh_ctx_create(int device, gsh_ctx** out) {
  GSH_API_BODY({
    GSH_REQUIRE(out, GSH_E_ARG, "out is NULL");
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    GSH_REQUIRE(device >= 0 && device < count, GSH_E_ARG, "no such device");
    auto c = std::make_unique<gsh_ctx>();
    c->device = device;
    c->set();
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming));
    // readback memory of the solve that allocates nothing (gsh_pcg_solve_static), and the
    // device address of its scratch array on this device
    c->pinned_slots(1);
    HIPCHK(hipHostMalloc((void**)&c->pinned_flag, 256, hipHostMallocDefault));
    static_scratch_setup(device);
    *out = c.release();
  })
}

static void ctx_free(gsh_ctx* ctx) {
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->pinned_flag) (void)hipHostFree(ctx->pinned_flag);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

static void ctx_retain(gsh_ctx* ctx) {
  if (ctx) ctx->refs.fetch_add(1);
}

static void ctx_release(gsh_ctx* ctx) {
  if (ctx && ctx->refs.fetch_sub(1) == 1 && ctx->closed.load()) ctx_free(ctx);
}

int gsh_ctx_destroy(gsh_ctx* ctx) {
  if (!ctx) return GSH_OK;
  ctx->closed.store(true);
  if (ctx->refs.load() == 0) {
    ctx_free(ctx);
  } else {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  return GSH_OK;
}

int gsh_ctx_mem_info(gsh_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes) {
  GSH_API_BODY({
    GSH_REQUIRE(ctx, GSH_E_ARG, "ctx is NULL");
    ctx->set();
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
  })
}

int gsh_ctx_synchronize(gsh_ctx* ctx) {
  GSH_API_BODY({
    GSH_REQUIRE(ctx, GSH_E_ARG, "ctx is NULL");
    ctx->set();
    HIPCHK(hipStreamSynchronize(ctx->stream));
  })
}

void gsh_step_params_default(gsh_step_params* p) {
  if (!p) return;
  // defaults of readNamelist.py:98-106, 181-193 and constants_template.h
  p->dt = 150.0;
  p->theta = 0.6;
  p->h0 = 0.01;
  p->ramp = 1.0;
  p->dtb_min = 10.0;
  p->dtb_max = 30.0;
  p->velmin_btrack = 1e-4;
  p->ihhat = 1;
  p->iettype = 2;
  p->rmaxvel = 5.0;  // readNamelist.py:192
  p->dfv0 = 0.01;    // readNamelist.py:130
  p->time = 0.0;
  p->shapiro0 = 0.5;  // readNamelist.py:95
  p->ishapiro = 1;    // readNamelist.py:94
}

static StepArgs step_args(const gsh_step_params* p) {
  GSH_REQUIRE(p, GSH_E_ARG, "step params are NULL");
  GSH_REQUIRE(p->dt > 0.0 && p->dtb_min > 0.0 && p->dtb_max > 0.0, GSH_E_ARG, "dt, dtb_min, dtb_max must be > 0");
  GSH_REQUIRE(p->theta >= 0.0 && p->theta <= 1.0, GSH_E_ARG, "theta must be in [0, 1]");
  GSH_REQUIRE(p->iettype >= 0 && p->iettype <= 5, GSH_E_ARG, "iettype must be in [0, 5]");
  GSH_REQUIRE(p->ishapiro == 0 || p->ishapiro == 1, GSH_E_ARG, "ishapiro must be 0 or 1");
  GSH_REQUIRE(!p->ishapiro || (p->shapiro0 >= 0.0 && p->shapiro0 <= 0.5), GSH_E_ARG,
              "shapiro0 must be in [0, 0.5]");
  StepArgs a{};
  a.dt = p->dt;
  a.theta = p->theta;
  a.h0 = p->h0;
  a.ramp = p->ramp;
  a.dtb_min = p->dtb_min;
  a.dtb_max = p->dtb_max;
  a.velmin2 = p->velmin_btrack * p->velmin_btrack;
  a.dfv0 = p->dfv0;
  a.rmaxvel = p->rmaxvel;
  a.time = p->time;
  a.shapiro0 = p->shapiro0;
  a.ishapiro = p->ishapiro;
  a.ihhat = p->ihhat;
  a.iettype = p->iettype;
  // backtrackAllNodes (meshdata.cu:3550-3551)
  a.ndelt_max = (int)std::fmax(1.0, p->dt / p->dtb_min);
  a.ndelt_min = (int)std::fmax(1.0, p->dt / p->dtb_max);
  return a;
}

int gsh_mesh_create(gsh_ctx* ctx, const gsh_mesh_desc* desc, gsh_mesh** out, int32_t* node_perm_out,
                    int32_t* elem_perm_out) {
  GSH_API_BODY({
    GSH_REQUIRE(desc && out, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(desc->n >= 3 && desc->ne >= 1, GSH_E_ARG, "mesh needs at least 3 nodes and 1 element");
    GSH_REQUIRE(desc->nvrt >= 2 && desc->nvrt <= 1024, GSH_E_ARG, "nvrt must be in [2, 1024]");
    GSH_REQUIRE(desc->xy && desc->elnode && desc->depth && (desc->sigma || desc->vgrid), GSH_E_ARG,
                "xy, elnode, depth and sigma (or a vertical grid) required");
    if (ctx) ctx->set();
    auto m = std::make_unique<gsh_mesh>();
    m->ctx = ctx;
    HostMesh& hm = m->hm;
    hm.n = desc->n;
    hm.ne = desc->ne;
    hm.nvrt = desc->nvrt;
    hm.x.resize(hm.n);
    hm.y.resize(hm.n);
    for (int32_t i = 0; i < hm.n; ++i) hm.x[i] = desc->xy[2 * i], hm.y[i] = desc->xy[2 * i + 1];
    hm.depth.assign(desc->depth, desc->depth + hm.n);
    if (desc->sigma)
      hm.sigma.assign(desc->sigma, desc->sigma + hm.nvrt);
    else
      hm.sigma.assign(hm.nvrt, 0.0);  // unused: the vertical grid gives the levels
    hm.elnode.assign(desc->elnode, desc->elnode + 3 * (size_t)hm.ne);
    hm.open_bnd.assign(hm.n, 0);
    if (desc->open_bnd)
      for (int32_t i = 0; i < hm.n; ++i) hm.open_bnd[i] = desc->open_bnd[i] ? 1 : 0;
    hm.kb.assign(hm.n, 0);
    if (desc->kb) hm.kb.assign(desc->kb, desc->kb + hm.n);
    for (int32_t i = 0; i < hm.n; ++i)
      GSH_REQUIRE(hm.kb[i] >= 0 && hm.kb[i] <= hm.nvrt - 2, GSH_E_ARG, "kb must be in [0, nvrt-2]");
    if (desc->vgrid) {
      const gsh_vgrid& v = *desc->vgrid;
      setup_vertical(hm, VGridIn{v.ivcor, v.sigma_node, v.nz, v.h_s, v.h_c, v.theta_b, v.theta_f, v.ztot, v.sigma_s});
    } else {
      for (int k = 1; k < hm.nvrt; ++k) GSH_REQUIRE(hm.sigma[k] > hm.sigma[k - 1], GSH_E_ARG, "sigma must increase");
    }
    std::vector<int32_t> nperm(hm.n), eperm(hm.ne);
    for (int32_t i = 0; i < hm.n; ++i) nperm[i] = i;
    for (int32_t i = 0; i < hm.ne; ++i) eperm[i] = i;
    if (desc->renumber) {
      for (int32_t k = 0; k < 3 * hm.ne; ++k)
        GSH_REQUIRE(hm.elnode[k] >= 0 && hm.elnode[k] < hm.n, GSH_E_TOPOLOGY, "element with an invalid node id");
      nperm = rcm_order(hm.n, hm.ne, hm.elnode);
      renumber(hm, nperm, eperm);
    }
    build_topology(hm, m->topo);
    build_geometry(hm, m->topo, m->geo);
    const Topology& t = m->topo;
    const int32_t n = hm.n, ne = hm.ne, ns = t.ns, nv = hm.nvrt;
    m->n_open = 0;
    for (auto v : hm.open_bnd) m->n_open += v;
    // ball -> CSR slots of the two other nodes of each ball element
    std::vector<int16_t> ball_slot(2 * t.ball_elem.size());
    for (int32_t i = 0; i < n; ++i) {
      const int32_t row = t.indptr[i], len = t.indptr[i + 1] - row;
      for (int32_t j = t.ball_ptr[i]; j < t.ball_ptr[i + 1]; ++j) {
        const int32_t e = t.ball_elem[j];
        const int id1 = t.ball_loc[j];
        for (int q = 0; q < 2; ++q) {
          const int32_t other = hm.elnode[3 * e + (id1 + 1 + q) % 3];
          int16_t slot = -1;
          for (int32_t k = 0; k < len; ++k)
            if (t.indices[row + k] == other) slot = (int16_t)k;
          GSH_REQUIRE(slot >= 0, GSH_E_TOPOLOGY, "ball element node missing from CSR row");
          ball_slot[2 * j + q] = slot;
        }
      }
    }
    if (node_perm_out) std::memcpy(node_perm_out, nperm.data(), sizeof(int32_t) * n);
    if (elem_perm_out) std::memcpy(elem_perm_out, eperm.data(), sizeof(int32_t) * ne);
    if (!ctx) {  // host-only mesh: topology and geometry without device state
      *out = m.release();
      return GSH_OK;
    }
    std::vector<int4> erec(2 * (size_t)ne);
    for (int32_t e = 0; e < ne; ++e) {
      erec[2 * e] = make_int4(hm.elnode[3 * e], hm.elnode[3 * e + 1], hm.elnode[3 * e + 2], t.elside[3 * e]);
      erec[2 * e + 1] = make_int4(t.elside[3 * e + 1], t.elside[3 * e + 2], 0, 0);
    }
    std::vector<int4> srec(ns);
    for (int32_t s = 0; s < ns; ++s)
      srec[s] = make_int4(t.isidenode[2 * s], t.isidenode[2 * s + 1], t.isdel[2 * s], t.isdel[2 * s + 1]);
    DevPool& P = m->pool;
    DMesh& d = m->d;
    d.n = n;
    d.ne = ne;
    d.ns = ns;
    d.nvrt = nv;
    d.x = P.upload(hm.x);
    d.y = P.upload(hm.y);
    d.depth = P.upload(hm.depth);
    d.kb = P.upload(hm.kb);
    d.sigma = P.upload(hm.sigma);
    d.vmode = hm.vmode;
    if (hm.vmode == 1) d.sig_node = P.upload(hm.sig_node);
    if (hm.vmode == 2) {
      d.ztot = P.upload(hm.ztot);
      d.ssig = P.upload(hm.ssig);
      d.cs = P.upload(hm.cs);
      d.izlev = hm.nz - 1;
      d.h_s = hm.h_s;
      d.h_c = hm.h_c;
    }
    d.open_bnd = P.upload(hm.open_bnd);
    d.node_bside = reinterpret_cast<const int2*>(P.upload(t.node_bside));
    d.eta2 = P.alloc<double>(n);
    d.eta1 = P.alloc<double>(n);
    d.grav2 = P.upload(std::vector<double>(n, 9.81));
    d.cd = P.upload(std::vector<double>(n, 2.5e-3));
    d.eth = P.alloc<double>(n);
    d.elbc = P.alloc<double>(n);
    d.uvw = P.alloc<double>(4 * (int64_t)n * nv);
    d.nrec = P.alloc<double2>(2 * (int64_t)n);
    d.eta_lev = P.alloc<double>(n);
    d.node_wet = P.alloc<int8_t>(n);
    d.etype = P.upload(std::vector<int8_t>(n, -1));
    d.elem = P.upload(erec);
    d.area = P.upload(m->geo.area);
    d.gradphi = P.upload(m->geo.gradphi);
    d.elem_wet = P.alloc<int8_t>(ne);
    d.elem_kb = P.alloc<int32_t>(ne);
    d.econ = P.alloc<double>(24 * (int64_t)ne);
    d.ghat1 = P.alloc<double>(2 * (int64_t)ne);
    d.vsource = P.alloc<double>(ne);
    d.side = P.upload(srec);
    d.side_type = P.upload(t.side_type);
    d.side_nbr = P.upload(t.side_nbr);
    d.side_cx = P.upload(m->geo.side_cx);
    d.side_cy = P.upload(m->geo.side_cy);
    d.side_len = P.upload(m->geo.side_len);
    d.side_kb = P.alloc<int32_t>(ns);
    d.side_wet = P.alloc<int8_t>(ns);
    d.zc = P.alloc<double>((int64_t)ns * nv);
    d.su2 = P.alloc<double>(2 * (int64_t)ns * nv);
    d.btu = P.alloc<double>(2 * (int64_t)ns * nv);
    d.foot = P.alloc<int32_t>((int64_t)ns * nv);
    d.bstatus = P.alloc<int32_t>((int64_t)ns * nv);
    d.sU = P.alloc<double>(2 * (int64_t)ns);
    d.side_us = P.alloc<double>(2 * (int64_t)ns);
    d.hhat = P.alloc<double>(ns);
    d.chi = P.alloc<double>(ns);
    d.ball_ptr = P.upload(t.ball_ptr);
    d.ball_elem = P.upload(t.ball_elem);
    d.ball_loc = P.upload(t.ball_loc);
    d.ball_slot = P.upload(ball_slot);
    d.nbr_ptr = P.upload(t.nbr_ptr);
    d.nbr_node = P.upload(t.nbr_node);
    d.csr_slot = P.upload(t.csr_slot);
    d.indptr = P.upload(t.indptr);
    d.diag_off = P.upload(t.diag_off);
    // initial levels with eta2 = 0
    gsh_step_params p;
    gsh_step_params_default(&p);
    launch_levels(d, step_args(&p), ctx->stream);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx_retain(ctx);
    *out = m.release();
  })
}

int gsh_mesh_destroy(gsh_mesh* mesh) {
  if (!mesh) return GSH_OK;
  gsh_ctx* ctx = mesh->ctx;
  if (ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  delete mesh;
  ctx_release(ctx);
  return GSH_OK;
}

int gsh_mesh_get_info(const gsh_mesh* m, gsh_mesh_info* out) {
  GSH_API_BODY({
    GSH_REQUIRE(m && out, GSH_E_ARG, "NULL argument");
    out->n = m->hm.n;
    out->ne = m->hm.ne;
    out->ns = m->topo.ns;
    out->nvrt = m->hm.nvrt;
    out->nnz = m->topo.indptr.back();
    out->max_row = m->topo.max_row;
    out->n_open = m->n_open;
  })
}

int gsh_mesh_get_topology(const gsh_mesh* m, int what, void* out) {
  GSH_API_BODY({
    GSH_REQUIRE(m && out, GSH_E_ARG, "NULL argument");
    const Topology& t = m->topo;
    auto put = [&](const auto& v) { std::memcpy(out, v.data(), v.size() * sizeof(v[0])); };
    switch (what) {
      case GSH_T_XY: {
        double* o = static_cast<double*>(out);
        for (int32_t i = 0; i < m->hm.n; ++i) o[2 * i] = m->hm.x[i], o[2 * i + 1] = m->hm.y[i];
        break;
      }
      case GSH_T_ELNODE: put(m->hm.elnode); break;
      case GSH_T_ELSIDE: put(t.elside); break;
      case GSH_T_ISDEL: put(t.isdel); break;
      case GSH_T_ISIDENODE: put(t.isidenode); break;
      case GSH_T_BALL_PTR: put(t.ball_ptr); break;
      case GSH_T_BALL_ELEM: put(t.ball_elem); break;
      case GSH_T_NBR_NODE: put(t.nbr_node); break;
      case GSH_T_CSR_INDPTR: put(t.indptr); break;
      case GSH_T_CSR_INDICES: put(t.indices); break;
      case GSH_T_CSR_SLOT: put(t.csr_slot); break;
      case GSH_T_DEPTH: put(m->hm.depth); break;
      case GSH_T_OPEN_BND: put(m->hm.open_bnd); break;
      case GSH_T_KB: put(m->hm.kb); break;
      case GSH_T_SIDE_TYPE: put(t.side_type); break;
      default: throw Error(GSH_E_ARG, "unknown topology array");
    }
  })
}

struct FieldRef {
  void* ptr;
  int64_t count;  // public element count
  int elem;       // bytes per public element
};

static FieldRef field_ref(const gsh_mesh* m, int field) {
  const DMesh& d = m->d;
  const int64_t n = d.n, ne = d.ne, ns = d.ns, nv = d.nvrt;
  switch (field) {
    case GSH_F_ETA2: return {d.eta2, n, 8};
    case GSH_F_ETA1: return {d.eta1, n, 8};
    case GSH_F_NODE_UVW: return {d.uvw, 3 * n * nv, 8};
    case GSH_F_SIDE_U2: return {d.su2, 2 * ns * nv, 8};
    case GSH_F_BTU: return {d.btu, 2 * ns * nv, 8};
    case GSH_F_FOOT: return {d.foot, ns * nv, 4};
    case GSH_F_BSTATUS: return {d.bstatus, ns * nv, 4};
    case GSH_F_CD: return {d.cd, n, 8};
    case GSH_F_GRAV2: return {d.grav2, n, 8};
    case GSH_F_ETH: return {d.eth, n, 8};
    case GSH_F_VSOURCE: return {d.vsource, ne, 8};
    case GSH_F_SIDE_U: return {d.sU, 2 * ns, 8};
    case GSH_F_SIDE_HHAT: return {d.hhat, ns, 8};
    case GSH_F_SIDE_CHI: return {d.chi, ns, 8};
    case GSH_F_GHAT1: return {d.ghat1, 2 * ne, 8};
    case GSH_F_ZC: return {d.zc, ns * nv, 8};
    case GSH_F_ELBC: return {d.elbc, n, 8};
    case GSH_F_WET: return {nullptr, n + ne + ns, 1};
    case GSH_F_DEPTH: return {const_cast<double*>(d.depth), n, 8};
    case GSH_F_ELEM_W: return {d.ew, ne * nv, 8};
    case GSH_F_ETYPE: return {d.etype, n, 1};
    default: throw Error(GSH_E_ARG, "unknown field");
  }
}

// The element vertical velocity is only read by gsh_nodal_velocity; its buffer (zero) is taken
// from the mesh's pool when it is first set, exported or used.
static void ensure_elem_w(gsh_mesh* m) {
  GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
  if (!m->d.ew) m->d.ew = m->pool.alloc<double>((int64_t)m->d.ne * m->d.nvrt);
}

int gsh_mesh_set_field(gsh_mesh* m, int field, const void* host, int64_t count) {
  GSH_API_BODY({
    GSH_REQUIRE(m && host, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    GSH_REQUIRE(field != GSH_F_WET, GSH_E_ARG, "wet flags are derived by gsh_update_levels");
    if (field == GSH_F_ELEM_W) ensure_elem_w(m);
    FieldRef f = field_ref(m, field);
    GSH_REQUIRE(count == f.count, GSH_E_ARG, "field size mismatch");
    if (field == GSH_F_ETYPE) {  // -1 (the step parameter) or an elevation BC type 0-5
      const int8_t* t = static_cast<const int8_t*>(host);
      for (int64_t i = 0; i < count; ++i)
        GSH_REQUIRE(t[i] >= -1 && t[i] <= 5, GSH_E_ARG, "elevation BC types must be in [-1, 5]");
    }
    std::vector<int32_t> sz_kb;
    if (field == GSH_F_DEPTH && m->hm.vmode == 2) {
      // SZ: the bottom levels follow the depth (Node::updateKb); a depth below the deepest Z
      // level is refused before anything changes
      HostMesh tmp = m->hm;
      tmp.depth.assign(static_cast<const double*>(host), static_cast<const double*>(host) + count);
      sz_bottom_levels(tmp);
      sz_kb = std::move(tmp.kb);
    }
    hipStream_t st = m->ctx->stream;
    if (field == GSH_F_NODE_UVW) {
      StreamBuffers bufs(st);
      double* tmp = bufs.alloc<double>((size_t)count);
      HIPCHK(hipMemcpyAsync(tmp, host, sizeof(double) * count, hipMemcpyHostToDevice, st));
      const int64_t recs = count / 3;
      hipLaunchKernelGGL(k_pack_uvw, dim3((unsigned)((recs + 255) / 256)), dim3(256), 0, st, tmp, m->d.uvw, recs);
    } else {
      HIPCHK(hipMemcpyAsync(f.ptr, host, (size_t)count * f.elem, hipMemcpyHostToDevice, st));
    }
    if (field == GSH_F_DEPTH) {  // the host copy answers GSH_T_DEPTH
      const double* h = static_cast<const double*>(host);
      if (m->hm.vmode == 2) {
        m->hm.kb = sz_kb;
        HIPCHK(hipMemcpyAsync(const_cast<int32_t*>(m->d.kb), m->hm.kb.data(), sizeof(int32_t) * m->hm.n,
                              hipMemcpyHostToDevice, st));
      }
      m->hm.depth.assign(h, h + count);
    }
    HIPCHK(hipStreamSynchronize(st));
  })
}

int gsh_mesh_get_field(const gsh_mesh* m, int field, void* host, int64_t count) {
  GSH_API_BODY({
    GSH_REQUIRE(m && host, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    FieldRef f = field_ref(m, field);
    GSH_REQUIRE(count == f.count, GSH_E_ARG, "field size mismatch");
    hipStream_t st = m->ctx->stream;
    if (field == GSH_F_ELEM_W && !f.ptr) {  // never set: zero
      std::memset(host, 0, (size_t)count * f.elem);
      return GSH_OK;
    }
    if (field == GSH_F_NODE_UVW) {
      StreamBuffers bufs(st);
      double* tmp = bufs.alloc<double>((size_t)count);
      const int64_t recs = count / 3;
      hipLaunchKernelGGL(k_unpack_uvw, dim3((unsigned)((recs + 255) / 256)), dim3(256), 0, st, m->d.uvw, tmp, recs);
      HIPCHK(hipMemcpyAsync(host, tmp, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    } else if (field == GSH_F_WET) {
      int8_t* h = static_cast<int8_t*>(host);
      HIPCHK(hipMemcpyAsync(h, m->d.node_wet, m->d.n, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(h + m->d.n, m->d.elem_wet, m->d.ne, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(h + m->d.n + m->d.ne, m->d.side_wet, m->d.ns, hipMemcpyDeviceToHost, st));
    } else {
      HIPCHK(hipMemcpyAsync(host, f.ptr, (size_t)count * f.elem, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
  })
}

int gsh_mesh_field_ptr(gsh_mesh* m, int field, void** dev_ptr, int64_t* count) {
  GSH_API_BODY({
    GSH_REQUIRE(m && dev_ptr, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(field != GSH_F_WET && field != GSH_F_NODE_UVW, GSH_E_ARG,
                "this field has a device layout of its own; use gsh_mesh_get_field");
    if (field == GSH_F_ELEM_W) ensure_elem_w(m);
    FieldRef f = field_ref(m, field);
    *dev_ptr = f.ptr;
    if (count) *count = f.count;
  })
}

int gsh_update_levels(gsh_mesh* m, const gsh_step_params* p) {
  GSH_API_BODY({
    GSH_REQUIRE(m, GSH_E_ARG, "mesh is NULL");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    launch_levels(m->d, step_args(p), m->ctx->stream);
    HIPCHK(hipGetLastError());
  })
}

int gsh_mesh_set_backtrack_staging(gsh_mesh* m, int32_t staged) {
  GSH_API_BODY({
    GSH_REQUIRE(m, GSH_E_ARG, "mesh is NULL");
    GSH_REQUIRE(staged == 0 || staged == 1, GSH_E_ARG, "staged must be 0 or 1");
    m->d.bt_stage = staged;
  })
}

int gsh_backtrack(gsh_mesh* m, const gsh_step_params* p, double* btu_host, int32_t* foot_host) {
  GSH_API_BODY({
    GSH_REQUIRE(m, GSH_E_ARG, "mesh is NULL");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    hipStream_t st = m->ctx->stream;
    launch_backtrack(m->d, step_args(p), nullptr, nullptr, 0, st);
    HIPCHK(hipGetLastError());
    const int64_t cnt = (int64_t)m->d.ns * m->d.nvrt;
    if (btu_host) HIPCHK(hipMemcpyAsync(btu_host, m->d.btu, 16 * cnt, hipMemcpyDeviceToHost, st));
    if (foot_host) HIPCHK(hipMemcpyAsync(foot_host, m->d.foot, 4 * cnt, hipMemcpyDeviceToHost, st));
    if (btu_host || foot_host) HIPCHK(hipStreamSynchronize(st));
  })
}

int gsh_backtrack_list(gsh_mesh* m, const gsh_step_params* p, const int32_t* sides, const int32_t* levels,
                       int32_t count) {
  GSH_API_BODY({
    GSH_REQUIRE(m && sides && levels && count >= 0, GSH_E_ARG, "bad arguments");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    for (int32_t i = 0; i < count; ++i)
      GSH_REQUIRE(sides[i] >= 0 && sides[i] < m->d.ns && levels[i] >= 0 && levels[i] < m->d.nvrt, GSH_E_ARG,
                  "start point out of range");
    if (count == 0) return GSH_OK;
    hipStream_t st = m->ctx->stream;
    StreamBuffers bufs(st);
    int32_t* buf = bufs.alloc<int32_t>(2 * (size_t)count);
    HIPCHK(hipMemcpyAsync(buf, sides, 4 * (size_t)count, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(buf + count, levels, 4 * (size_t)count, hipMemcpyHostToDevice, st));
    launch_backtrack(m->d, step_args(p), buf, buf + count, count, st);
    HIPCHK(hipStreamSynchronize(st));
  })
}

int gsh_nodal_velocity(gsh_mesh* m, const gsh_step_params* p) {
  GSH_API_BODY({
    GSH_REQUIRE(m && p, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    GSH_REQUIRE(p->rmaxvel >= 0.0, GSH_E_ARG, "rmaxvel must be >= 0");
    m->ctx->set();
    ensure_elem_w(m);
    launch_nodal_velocity(m->d, p->rmaxvel, m->ctx->stream);
    HIPCHK(hipGetLastError());
  })
}

int gsh_momentum(gsh_mesh* m, const gsh_step_params* p) {
  GSH_API_BODY({
    GSH_REQUIRE(m && p, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    GSH_REQUIRE(p->rmaxvel >= 0.0, GSH_E_ARG, "rmaxvel must be >= 0");
    GSH_REQUIRE(p->dfv0 >= 0.0, GSH_E_ARG, "dfv0 must be >= 0");
    GSH_REQUIRE(p->ishapiro == 0 || p->ishapiro == 1, GSH_E_ARG, "ishapiro must be 0 or 1");
    const StepArgs a = step_args(p);
    m->ctx->set();
    launch_momentum(m->d, a, m->ctx->stream);
    if (a.ishapiro == 1) {
      // the filtered field is written whole into a second [ns][nvrt][2] buffer of the mesh's
      // pool, which then becomes u2 (the reference filters through its baroclinic array bcc and
      // copies back); device pointers to GSH_F_SIDE_U2 taken before this call go stale
      if (!m->shap_tmp) m->shap_tmp = m->pool.alloc<double>(2 * (int64_t)m->d.ns * m->d.nvrt);
      launch_shapiro(m->d, a.shapiro0, m->shap_tmp, m->ctx->stream);
      std::swap(m->d.su2, m->shap_tmp);
    }
    HIPCHK(hipGetLastError());
  })
}

int gsh_set_tides(gsh_mesh* m, int32_t nbfr, const double* amig, const double* ff, const double* face,
                  const double* emo, const double* efa) {
  GSH_API_BODY({
    GSH_REQUIRE(m, GSH_E_ARG, "mesh is NULL");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    GSH_REQUIRE(nbfr >= 0 && nbfr <= 256, GSH_E_ARG, "nbfr must be in [0, 256]");
    GSH_REQUIRE(nbfr == 0 || (amig && ff && face && emo && efa), GSH_E_ARG, "NULL tidal array");
    m->ctx->set();
    DMesh& d = m->d;
    if (nbfr == 0) {
      d.nbfr = 0;
      return GSH_OK;
    }
    const int64_t per = (int64_t)d.n * nbfr;
    if (per > m->tide_cap) {  // the pool keeps its chunks; a larger table takes a new one
      m->tide_emo = m->pool.alloc<double>(per);
      m->tide_efa = m->pool.alloc<double>(per);
      m->tide_cap = per;
    }
    if (nbfr > m->tide_nb) {
      m->tide_glob = m->pool.alloc<double>(3 * (int64_t)nbfr);
      m->tide_nb = nbfr;
    }
    std::vector<double> g(3 * (size_t)nbfr);
    for (int j = 0; j < nbfr; ++j) {
      g[j] = amig[j];
      g[nbfr + j] = ff[j];
      g[2 * nbfr + j] = face[j];
    }
    hipStream_t st = m->ctx->stream;
    HIPCHK(hipMemcpyAsync(m->tide_glob, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(m->tide_emo, emo, sizeof(double) * per, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(m->tide_efa, efa, sizeof(double) * per, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    d.nbfr = nbfr;
    d.tide = m->tide_glob;
    d.emo = m->tide_emo;
    d.efa = m->tide_efa;
  })
}

int gsh_vertical_velocity(gsh_mesh* m, const gsh_step_params* p) {
  GSH_API_BODY({
    GSH_REQUIRE(m && p, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    ensure_elem_w(m);
    launch_vertical_velocity(m->d, p->h0, m->ctx->stream);
    HIPCHK(hipGetLastError());
  })
}

int gsh_rhs_producers(gsh_mesh* m, const gsh_step_params* p) {
  GSH_API_BODY({
    GSH_REQUIRE(m, GSH_E_ARG, "mesh is NULL");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    launch_rhs(m->d, step_args(p), m->ctx->stream);
    HIPCHK(hipGetLastError());
  })
}

// ---------------------------------------------------------------------- solver entry points
static std::vector<int32_t> diag_abs_of(const std::vector<int32_t>& indptr, const std::vector<int32_t>& indices,
                                        int32_t n) {
  std::vector<int32_t> d(n, -1);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k)
      if (indices[k] == i) d[i] = k;
  for (int32_t i = 0; i < n; ++i)
    if (d[i] < 0) throw Error(GSH_E_ARG, "CSR row " + std::to_string(i) + " has no diagonal entry");
  return d;
}

// A solver on the mesh's pattern; an SSOR solver gets its layout for `ssor_rpt` rows per thread.
static gsh_solver* mesh_solver_new(gsh_mesh* mesh, int precision, int precond, int ssor_rpt) {
  GSH_REQUIRE(precision == GSH_PREC_FP64 || precision == GSH_PREC_MIXED, GSH_E_ARG, "unknown precision");
  GSH_REQUIRE(precond == GSH_PRECOND_JACOBI || precond == GSH_PRECOND_SSOR, GSH_E_ARG, "unknown preconditioner");
  GSH_REQUIRE(mesh->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
  mesh->ctx->set();
  auto s = std::make_unique<gsh_solver>();
  s->ctx = mesh->ctx;
  s->mesh = mesh;
  s->precision = precision;
  s->precond = precond;
  const Topology& t = mesh->topo;
  std::vector<int32_t> dabs(mesh->hm.n);
  for (int32_t i = 0; i < mesh->hm.n; ++i) dabs[i] = t.indptr[i] + t.diag_off[i];
  solver_alloc(s.get(), mesh->hm.n, t.indptr.back(), t.indptr, t.indices, dabs);
  if (precond == GSH_PRECOND_SSOR) {
    s->rpt = ssor_rpt;
    s->ntiles = s->tiles_for(s->n);
    s->hs.ntiles = s->ntiles;
    s->h_indptr = t.indptr;
    s->h_indices = t.indices;
    ssor_build(s.get());
  }
  ctx_retain(s->ctx);
  return s.release();
}

int gsh_solver_create(gsh_mesh* mesh, int precision, int precond, gsh_solver** out) {
  GSH_API_BODY({
    GSH_REQUIRE(mesh && out, GSH_E_ARG, "NULL argument");
    *out = mesh_solver_new(mesh, precision, precond, 1);
  })
}

int gsh_solver_create_csr(gsh_ctx* ctx, int32_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                          int precision, int precond, gsh_solver** out) {
  GSH_API_BODY({
    GSH_REQUIRE(ctx && indptr && indices && out && n > 0 && nnz > 0, GSH_E_ARG, "bad arguments");
    GSH_REQUIRE(precision == GSH_PREC_FP64 || precision == GSH_PREC_MIXED, GSH_E_ARG, "unknown precision");
    GSH_REQUIRE(precond == GSH_PRECOND_JACOBI || precond == GSH_PRECOND_SSOR, GSH_E_ARG, "unknown preconditioner");
    GSH_REQUIRE(nnz < (int64_t)1 << 31, GSH_E_ARG, "nnz exceeds int32 indexing");
    GSH_REQUIRE(indptr[0] == 0 && indptr[n] == nnz, GSH_E_ARG, "indptr must start at 0 and end at nnz");
    for (int32_t i = 0; i < n; ++i) GSH_REQUIRE(indptr[i + 1] >= indptr[i], GSH_E_ARG, "indptr must be non-decreasing");
    for (int64_t k = 0; k < nnz; ++k) GSH_REQUIRE(indices[k] >= 0 && indices[k] < n, GSH_E_ARG, "column out of range");
    ctx->set();
    auto s = std::make_unique<gsh_solver>();
    s->ctx = ctx;
    s->precision = precision;
    s->precond = precond;
    std::vector<int32_t> ip(indptr, indptr + n + 1), ix(indices, indices + nnz);
    solver_alloc(s.get(), n, nnz, ip, ix, diag_abs_of(ip, ix, n));
    if (precond == GSH_PRECOND_SSOR) {
      s->rpt = 1;
      s->ntiles = s->tiles_for(s->n);
      s->hs.ntiles = s->ntiles;
      s->h_indptr = std::move(ip);
      s->h_indices = std::move(ix);
      ssor_build(s.get());
    }
    ctx_retain(s->ctx);
    *out = s.release();
  })
}

int gsh_solver_destroy(gsh_solver* s) {
  if (!s) return GSH_OK;
  gsh_ctx* ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete s;
  ctx_release(ctx);
  return GSH_OK;
}

int gsh_solver_csr(gsh_solver* s, double** data, int32_t** indices, int32_t** indptr, int64_t* nnz) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    if (data) *data = s->val;
    if (indices) *indices = s->col;
    if (indptr) *indptr = s->rowptr;
    if (nnz) *nnz = s->nnz;
  })
}

int gsh_solver_set_values(gsh_solver* s, const double* data) {
  GSH_API_BODY({
    GSH_REQUIRE(s && data, GSH_E_ARG, "NULL argument");
    s->ctx->set();
    HIPCHK(hipMemcpy(s->val, data, sizeof(double) * s->nnz, hipMemcpyHostToDevice));
  })
}

int gsh_solver_get_values(gsh_solver* s, double* data) {
  GSH_API_BODY({
    GSH_REQUIRE(s && data, GSH_E_ARG, "NULL argument");
    s->ctx->set();
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    HIPCHK(hipMemcpy(data, s->val, sizeof(double) * s->nnz, hipMemcpyDeviceToHost));
  })
}

int gsh_solver_vec_ptr(gsh_solver* s, double** b, double** x) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    if (b) *b = s->b;
    if (x) *x = s->x;
  })
}

int gsh_solver_reset_history(gsh_solver* s) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    s->ctx->set();
    for (auto* h : s->h) launch_fill_f64(h, 0.0, s->n, s->ctx->stream);
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
  })
}

int gsh_solver_set_tuning(gsh_solver* s, int rows_per_thread, double inner_rtol) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    GSH_REQUIRE(rows_per_thread == 1 || rows_per_thread == 2 || rows_per_thread == 4 || rows_per_thread == 8,
                GSH_E_ARG, "rows_per_thread must be 1, 2, 4 or 8");
    GSH_REQUIRE(inner_rtol > 0.0 && inner_rtol < 1.0, GSH_E_ARG, "inner_rtol must be in (0, 1)");
    s->ctx->set();
    const bool relayout = s->ssor && s->ssor->rpt != rows_per_thread;
    s->rpt = rows_per_thread;
    s->ntiles = s->tiles_for(s->n);
    s->hs.ntiles = s->ntiles;
    s->hs32.ntiles = s->ntiles;
    s->inner_rtol = inner_rtol;
    if (relayout) {
      HIPCHK(hipStreamSynchronize(s->ctx->stream));
      ssor_build(s);
    }
  })
}

int gsh_solver_autotune(gsh_solver* s, int iters, int* best_rpt, double* us_per_iter) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    GSH_REQUIRE(iters >= 2, GSH_E_ARG, "iters must be >= 2");
    s->ctx->set();
    hipStream_t st = s->ctx->stream;
    const int cand[4] = {1, 2, 4, 8};
    double best = 1e300;
    int pick = s->rpt;
    for (int c = 0; c < 4; ++c) {
      const int r = cand[c];
      s->rpt = r;
      s->ntiles = s->tiles_for(s->n);
      s->hs.ntiles = s->ntiles;
      s->hs32.ntiles = s->ntiles;
      double us = 0.0;
      if (s->precond == GSH_PRECOND_SSOR) {
        if (!s->ssor || s->ssor->rpt != r) {
          HIPCHK(hipStreamSynchronize(st));
          ssor_build(s);
        }
        SsorTimes tt;
        if (s->precision == GSH_PREC_MIXED) {
          ssor_sync_desc32(s, s->b);
          run_ssor<float>(s, 1e-300, iters, iters, 60.0, &tt);
        } else {
          ssor_sync_desc(s);
          run_ssor<double>(s, 1e-300, iters, iters, 60.0, &tt);
        }
        us = tt.iters ? 1e3 * (tt.ms_bwd + tt.ms_fwd + tt.ms_upd) / tt.iters : 1e300;
      } else {
        solver_sync_desc(s);
        PcgOutcome o;
        PhaseTimes pt;
        run_pcg<double>(s->ctx, s->ds, 1, s->dtiles, s->ntiles, r, 1e-300, iters, iters, 60.0, o, nullptr, -1, &pt);
        us = pt.launches ? 1e3 * (pt.ms_a + pt.ms_b) / pt.launches : 1e300;
      }
      if (us_per_iter) us_per_iter[c] = us;
      if (us < best) {
        best = us;
        pick = r;
      }
    }
    HIPCHK(hipStreamSynchronize(st));
    s->rpt = pick;
    s->ntiles = s->tiles_for(s->n);
    s->hs.ntiles = s->ntiles;
    s->hs32.ntiles = s->ntiles;
    if (s->precond == GSH_PRECOND_SSOR && s->ssor->rpt != pick) ssor_build(s);
    if (best_rpt) *best_rpt = pick;
  })
}

int gsh_solver_set_graphs(gsh_solver* s, int on) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    s->ctx->set();
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    for (WindowGraph* g : {&s->wg_jacobi, &s->wg_ssor}) {
      g->reset();
      g->failed = false;
      g->enabled = on != 0;
    }
  })
}

int gsh_solver_set_omega(gsh_solver* s, double omega) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    GSH_REQUIRE(s->precond == GSH_PRECOND_SSOR, GSH_E_STATE, "omega applies to the SSOR preconditioner");
    GSH_REQUIRE(omega > 0.0 && omega < 2.0, GSH_E_ARG, "omega must be in (0, 2)");
    s->omega = omega;
  })
}

int gsh_solver_ssor_info(gsh_solver* s, int32_t* ncolours, int32_t* padded_rows, int32_t* colour_out) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    GSH_REQUIRE(s->ssor, GSH_E_STATE, "not an SSOR solver");
    const SsorLayout& L = s->ssor->L;
    if (ncolours) *ncolours = L.ncol;
    if (padded_rows) *padded_rows = L.np;
    if (colour_out) std::copy(L.colour.begin(), L.colour.end(), colour_out);
  })
}

int gsh_colour_csr(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* colour_out,
                   int32_t* ncolours) {
  GSH_API_BODY({
    GSH_REQUIRE(n >= 0 && indptr && indices && colour_out && ncolours, GSH_E_ARG, "bad arguments");
    for (int32_t i = 0; i < n; ++i)
      GSH_REQUIRE(indptr[i + 1] >= indptr[i], GSH_E_ARG, "indptr must be non-decreasing");
    for (int64_t k = 0; k < indptr[n]; ++k)
      GSH_REQUIRE(indices[k] >= 0 && indices[k] < n, GSH_E_ARG, "column out of range");
    *ncolours = colour_smallest_last(n, indptr, indices, colour_out);
  })
}

int gsh_assemble(gsh_mesh* m, gsh_solver* s, const gsh_step_params* p, double time) {
  (void)time;  // only bed deformation (imm != 0) depends on time (meshdata.cu:3244-3258)
  GSH_API_BODY({
    GSH_REQUIRE(m && s, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(s->mesh == m, GSH_E_ARG, "solver is not bound to this mesh");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    launch_assemble(m->d, step_args(p), s->val, s->b, m->ctx->stream);
    HIPCHK(hipGetLastError());
  })
}

int gsh_solver_solve(gsh_solver* s, const double* b_host, double* x_host, double rtol, int maxiter,
                     int check_every, double timeout_s, int* iters_out, int* status_out) {
  GSH_API_BODY({
    GSH_REQUIRE(s, GSH_E_ARG, "solver is NULL");
    GSH_REQUIRE(rtol >= 0.0, GSH_E_ARG, "rtol must be >= 0");
    if (maxiter <= 0) maxiter = 10 * s->n;  // pcg_solver.py:1082-1083
    s->ctx->set();
    hipStream_t st = s->ctx->stream;
    const auto t0 = Clock::now();
    if (b_host) HIPCHK(hipMemcpyAsync(s->b, b_host, sizeof(double) * s->n, hipMemcpyHostToDevice, st));
    solver_sync_desc(s);
    int iters = 0, status = 0;
    if (s->precision == GSH_PREC_MIXED) {
      solve_mixed(s, rtol, maxiter, check_every, timeout_s, &iters, &status);
    } else if (s->precond == GSH_PRECOND_SSOR) {
      ssor_sync_desc(s);
      const PcgState r = run_ssor<double>(s, rtol, maxiter, check_every, timeout_s, nullptr, true);
      iters = iters_of(r, maxiter);
      status = (r.done >= 0) ? 0 : GSH_W_MAXITER;
      s->last.iters = iters;
      s->last.outer = 1;
      s->last.rr = r.rr_last;
      s->last.atol2 = r.atol2;
    } else {
      PcgOutcome o;
      run_pcg<double>(s->ctx, s->ds, 1, s->dtiles, s->ntiles, s->rpt, rtol, maxiter, check_every, timeout_s, o,
                      nullptr, -1, nullptr, nullptr, &s->wg_jacobi);
      const PcgState& r = o.st[0];
      iters = iters_of(r, maxiter);
      status = (r.done >= 0) ? 0 : GSH_W_MAXITER;
      s->last.iters = iters;
      s->last.outer = 1;
      s->last.rr = r.rr_last;
      s->last.atol2 = r.atol2;
    }
    s->last.true_rel = -1.0;
    if (x_host) HIPCHK(hipMemcpyAsync(x_host, s->x, sizeof(double) * s->n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    solver_roll(s);
    s->last.ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    if (iters_out) *iters_out = iters;
    if (status_out) *status_out = status;
  })
}

int gsh_pcg_solve_static(gsh_ctx* ctx, int32_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                         const double* data, const double* b, double* x, double rtol, int maxiter, int check_every,
                         double timeout_s, int* iters_out, int* status_out) {
  GSH_API_BODY({
    GSH_REQUIRE(ctx && indptr && indices && data && b && x, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(n >= 1 && nnz >= 1, GSH_E_ARG, "n and nnz must be positive");
    GSH_REQUIRE(n <= GSH_STATIC_MAX_ROWS, GSH_E_ARG,
                "n = " + std::to_string(n) + " exceeds the static workspace (GSH_STATIC_MAX_ROWS = " +
                    std::to_string(GSH_STATIC_MAX_ROWS) + " rows)");
    GSH_REQUIRE(nnz < (int64_t)1 << 31, GSH_E_ARG, "nnz exceeds int32 indexing");
    GSH_REQUIRE(rtol >= 0.0, GSH_E_ARG, "rtol must be >= 0");
    GSH_REQUIRE(ctx->device >= 0 && ctx->device < kStaticMaxDevices && g_scratch_dev[ctx->device], GSH_E_STATE,
                "no static workspace on this device");
    GSH_REQUIRE(ctx->pinned && ctx->pinned_cap >= 2 && ctx->pinned_flag, GSH_E_STATE, "context without readback memory");
    if (maxiter <= 0) maxiter = 10 * n;  // pcg_solver.py:1082-1083
    if (timeout_s <= 0.0) timeout_s = 60.0;
    ctx->set();
    hipStream_t st = ctx->stream;
    char* base = g_scratch_dev[ctx->device];
    const int ntiles = (n + kTpb * kStaticRpt - 1) / (kTpb * kStaticRpt);
    PcgSys<double>* dsys = reinterpret_cast<PcgSys<double>*>(base + kScrSys);
    PcgState* dgather = reinterpret_cast<PcgState*>(base + kScrGather);
    int32_t* dflag = reinterpret_cast<int32_t*>(base + kScrFlag);
    TileRef* dtiles = reinterpret_cast<TileRef*>(base + kScrTiles);
    int32_t* ddiag = reinterpret_cast<int32_t*>(base + kScrDiag);
    // the pattern check and the diagonal positions, read back before any row is used
    hipLaunchKernelGGL(k_static_clear, dim3(1), dim3(1), 0, st, dflag);
    hipLaunchKernelGGL(k_static_csr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, indptr, indices, n, nnz, ddiag,
                       dflag, dtiles, ntiles);
    HIPCHK(hipMemcpyAsync(ctx->pinned_flag, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    wait_event(ctx->ev[0], timeout_s, "CSR check");
    const int32_t f = *ctx->pinned_flag;
    GSH_REQUIRE(!(f & 1), GSH_E_ARG, "indptr must start at 0, be non-decreasing and end at nnz");
    GSH_REQUIRE(!(f & 2), GSH_E_ARG, "column out of range");
    GSH_REQUIRE(!(f & 4), GSH_E_ARG, "a CSR row has no diagonal entry");
    PcgSys<double> h{};
    h.n = n;
    h.ntiles = ntiles;
    h.rowptr = indptr;
    h.col = indices;
    h.val = data;
    h.diag = ddiag;
    h.cdel = nullptr;
    h.b = b;
    h.x = x;
    double* v = reinterpret_cast<double*>(base + kScrVec);
    const size_t vs = kScrVecBytes / sizeof(double);
    h.r = v;
    h.z = v + vs;
    h.pa = v + 2 * vs;
    h.pb = v + 3 * vs;
    h.q = v + 4 * vs;
    h.invd = v + 5 * vs;
    h.h0 = h.h1 = h.h2 = nullptr;  // cold start: zero initial guess
    double* part = reinterpret_cast<double*>(base + kScrPart);
    const size_t ps = align256(kStaticTiles * sizeof(double)) / sizeof(double);
    h.part_rr = part;
    h.part_rz = part + ps;
    h.part_pq = part + 2 * ps;
    h.part_bb = part + 3 * ps;
    h.st = reinterpret_cast<PcgState*>(base + kScrState);
    hipLaunchKernelGGL(k_static_desc, dim3(1), dim3(1), 0, st, h, dsys);
    HIPCHK(hipGetLastError());
    run_pcg_core<double>(ctx, dsys, 1, dtiles, ntiles, kStaticRpt, rtol, maxiter, check_every, timeout_s, dgather,
                         nullptr, -1, nullptr, nullptr, nullptr);
    const PcgState r = ctx->pinned[0];
    if (iters_out) *iters_out = iters_of(r, maxiter);
    if (status_out) *status_out = (r.done >= 0) ? 0 : GSH_W_MAXITER;
  })
}

int gsh_solver_check(gsh_solver* s, double* rel_out) {
  GSH_API_BODY({
    GSH_REQUIRE(s && rel_out, GSH_E_ARG, "NULL argument");
    s->ctx->set();
    hipStream_t st = s->ctx->stream;
    // the last solution is the newest history entry after the roll
    solver_sync_desc(s);
    launch_true_residual(s->ds, s->dtiles, s->ntiles, s->rpt, s->h[0], nullptr, st);
    launch_reduce_pair(s->hs.part_rr, s->hs.part_bb, s->ntiles, s->red2, st);
    double h[2];
    HIPCHK(hipMemcpyAsync(h, s->red2, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const double rn = std::sqrt(h[0]), bn = std::sqrt(h[1]);
    *rel_out = (bn > 0.0) ? rn / bn : rn;
    s->last.true_rel = *rel_out;
  })
}

int gsh_solver_last_stats(gsh_solver* s, gsh_solve_stats* out) {
  GSH_API_BODY({
    GSH_REQUIRE(s && out, GSH_E_ARG, "NULL argument");
    *out = s->last;
  })
}

int gsh_update_elevation(gsh_mesh* m, gsh_solver* s) {
  GSH_API_BODY({
    GSH_REQUIRE(m && s && s->mesh == m, GSH_E_ARG, "solver is not bound to this mesh");
    GSH_REQUIRE(m->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    m->ctx->set();
    // after the solve the solution is the newest history entry (scatterXToMesh)
    launch_update_elevation(m->d, s->h[0], m->ctx->stream);
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
  })
}

// ---------------------------------------------------------------------- batching
int gsh_batch_solve(gsh_solver* const* solvers, int count, double rtol, int maxiter, int check_every,
                    double timeout_s, int* iters_out) {
  GSH_API_BODY({
    GSH_REQUIRE(solvers && count > 0, GSH_E_ARG, "bad arguments");
    gsh_ctx* ctx = solvers[0]->ctx;
    for (int i = 0; i < count; ++i) {
      GSH_REQUIRE(solvers[i] && solvers[i]->ctx == ctx, GSH_E_ARG, "all solvers must share one context");
      GSH_REQUIRE(solvers[i]->precision == solvers[0]->precision, GSH_E_ARG,
                  "all solvers of a batch need the same precision");
      GSH_REQUIRE(solvers[i]->precond == solvers[0]->precond, GSH_E_ARG,
                  "all solvers of a batch need the same preconditioner");
      GSH_REQUIRE(solvers[i]->precond != GSH_PRECOND_SSOR || solvers[i]->rpt == solvers[0]->rpt, GSH_E_ARG,
                  "batched SSOR solvers need the same rows per thread (gsh_solver_set_tuning)");
      for (int j = 0; j < i; ++j) GSH_REQUIRE(solvers[j] != solvers[i], GSH_E_ARG, "a solver appears twice");
    }
    ctx->set();
    hipStream_t st = ctx->stream;
    const int rpt = solvers[0]->rpt;
    int mi = maxiter;
    std::vector<PcgSys<double>> hs(count);
    std::vector<TileRef> tiles;
    for (int i = 0; i < count; ++i) {
      gsh_solver* s = solvers[i];
      s->rpt = rpt;
      s->ntiles = s->tiles_for(s->n);
      s->hs.x = s->x;
      s->hs.h0 = s->h[0];
      s->hs.h1 = s->h[1];
      s->hs.h2 = s->h[2];
      s->hs.ntiles = s->ntiles;
      hs[i] = s->hs;
      if (mi <= 0) mi = std::max(mi, 10 * s->n);
      for (int t = 0; t < s->ntiles; ++t) tiles.push_back({i, t});
    }
    if (maxiter <= 0) {
      maxiter = 0;
      for (int i = 0; i < count; ++i) maxiter = std::max(maxiter, 10 * solvers[i]->n);
    }
    if (solvers[0]->precision == GSH_PREC_MIXED) {
      for (int i = 0; i < count; ++i) solvers[i]->hs32.ntiles = solvers[i]->ntiles;
      std::vector<int> its, sts;
      run_mixed(ctx, solvers, count, rtol, maxiter, check_every, timeout_s, its, sts);
      for (int i = 0; i < count; ++i) {
        if (iters_out) iters_out[i] = its[i];
        solvers[i]->last.true_rel = -1.0;
        solver_roll(solvers[i]);
      }
      return GSH_OK;
    }
    if (solvers[0]->precond == GSH_PRECOND_SSOR) {
      for (int i = 0; i < count; ++i) ssor_sync_desc(solvers[i]);
      std::vector<PcgState> out;
      run_ssor_batch<double>(solvers, count, std::vector<double>(count, rtol), maxiter, check_every, timeout_s,
                             std::vector<char>(count, 1), out);
      for (int i = 0; i < count; ++i) {
        gsh_solver* s = solvers[i];
        const int it = iters_of(out[i], maxiter);
        if (iters_out) iters_out[i] = it;
        s->last.iters = it;
        s->last.outer = 1;
        s->last.rr = out[i].rr_last;
        s->last.atol2 = out[i].atol2;
        s->last.true_rel = -1.0;
        solver_roll(s);
      }
      return GSH_OK;
    }
    StreamBuffers tmp(st);
    PcgSys<double>* dsys = tmp.alloc<PcgSys<double>>(count);
    TileRef* dt = tmp.alloc<TileRef>(tiles.size());
    HIPCHK(hipMemcpyAsync(dsys, hs.data(), sizeof(PcgSys<double>) * count, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dt, tiles.data(), sizeof(TileRef) * tiles.size(), hipMemcpyHostToDevice, st));
    PcgOutcome o;
    run_pcg<double>(ctx, dsys, count, dt, (int)tiles.size(), rpt, rtol, maxiter, check_every, timeout_s, o);
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < count; ++i) {
      gsh_solver* s = solvers[i];
      const int it = iters_of(o.st[i], maxiter);
      if (iters_out) iters_out[i] = it;
      s->last.iters = it;
      s->last.outer = 1;
      s->last.rr = o.st[i].rr_last;
      s->last.atol2 = o.st[i].atol2;
      solver_roll(s);
    }
  })
}

}  // extern "C"

// ====================================================================== serving
// Continuous batching: requests are admitted between check windows as device memory and
// slots allow; all admitted systems advance together, one lock-step launch per PCG phase,
// each with its own iteration counter and stopping test; finished systems leave at the next
// window boundary. Per-system reductions use per-system tiles only, so a request's result is
// identical alone and in any mix.
namespace {
struct Request {
  int64_t ticket = 0;
  int32_t mesh_id = 0, n = 0;
  // page-locked host vectors of the server's per-mesh pools: the right-hand side from submit
  // until the request retires, the solution from retirement until gsh_server_release
  double* hb = nullptr;
  double* hx = nullptr;
  bool has_b = false, with_bt = false;
  Clock::time_point t_submit;
  int status = 1;  // 1 queued/active, 0 done, GSH_E_* failed
  bool active = false;
  int32_t iters = 0;
  int32_t run = 0;  // iterations issued for this request
  double latency_ms = 0.0;
  gsh_solver* solver = nullptr;
  int64_t bytes = 0;
  // mixed precision: outer refinement state
  int phase = 0;          // 0: FP64 residual check due, 1: FP32 inner solve running
  bool inner_fresh = false;
  bool ssor_gathered = false;  // SSOR inner solves: FP32 ELL values gathered by the first one
  int outer = 0, inner_total = 0;
  double atol2 = -1.0, in_rtol = 0.0;
  hipEvent_t ready = nullptr;  // end of the request's set-up on the staging stream
  ~Request() {
    if (ready) (void)hipEventDestroy(ready);
  }
};
}  // namespace

struct gsh_server {
  gsh_ctx* ctx = nullptr;
  gsh_server_config cfg{};
  int64_t budget = 0, used = 0;
  std::vector<gsh_mesh*> meshes;
  std::vector<gsh_step_params> params;           // step parameters of each mesh's admission set-up
  std::vector<std::vector<gsh_solver*>> pool;  // idle solvers per mesh
  // idle page-locked host vectors of each mesh's length: copies from and to them run
  // asynchronously at full link speed, where pageable copies block the host and fault in fresh
  // pages for every request
  std::vector<std::vector<double*>> hfree;
  std::vector<Request*> retiring;  // solves ended in this step, their x copies in flight
  std::deque<Request*> queue;
  std::vector<Request*> active;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_ticket = 1;
  PcgSys<double>* dsys = nullptr;
  int dsys_cap = 0;
  TileRef* dtiles = nullptr;
  int dtiles_cap = 0;
  TileRef* dfresh = nullptr;
  int dfresh_cap = 0;
  PcgState* dgather = nullptr;
  int dgather_cap = 0;
  PcgSys<float>* dsys32 = nullptr;
  int dsys32_cap = 0;
  double* dred = nullptr;
  int dred_cap = 0;
  SsorSysT<double>* dssor = nullptr;
  int dssor_cap = 0;
  SsorSysT<float>* dssor32 = nullptr;
  int dssor32_cap = 0;
  SsorRef* drefs = nullptr;
  int drefs_cap = 0;
  // admission set-up (backtracking, assembly, casts) runs on its own stream and overlaps the
  // solve window of the step that admitted the request; the request joins the next step
  hipStream_t stage = nullptr;
  std::vector<Request*> staged;
  std::deque<int64_t> finished;  // tickets retired and not yet reported by gsh_server_finished
  ~gsh_server() {
    for (auto& v : pool)
      for (auto* s : v) gsh_solver_destroy(s);
    for (auto& kv : reqs) {
      if (kv.second->solver) gsh_solver_destroy(kv.second->solver);
      if (kv.second->hb) (void)hipHostFree(kv.second->hb);
      if (kv.second->hx) (void)hipHostFree(kv.second->hx);
    }
    for (auto& v : hfree)
      for (double* p : v) (void)hipHostFree(p);
    (void)hipFree(dsys);
    (void)hipFree(dtiles);
    (void)hipFree(dfresh);
    (void)hipFree(dgather);
    (void)hipFree(dsys32);
    (void)hipFree(dred);
    (void)hipFree(dssor);
    (void)hipFree(dssor32);
    (void)hipFree(drefs);
    if (stage) (void)hipStreamDestroy(stage);
  }
};

template <class T>
static void ensure_cap(T*& p, int& cap, int need) {
  if (need <= cap) return;
  if (p) HIPCHK(hipFree(p));
  p = nullptr;
  const int c = std::max(need, 2 * cap);
  HIPCHK(hipMalloc((void**)&p, sizeof(T) * c));
  cap = c;
}

// Rows per thread of the server's solvers.
static int server_rpt(const gsh_server_config& cfg) { return cfg.precond == GSH_PRECOND_SSOR ? 2 : 4; }

static int64_t solver_bytes(const gsh_mesh* m, int precision, int precond) {
  const int64_t n = m->hm.n, nnz = m->topo.indptr.back();
  int64_t b = 14 * nnz + 4 * (n + 1) + 4 * n + 8 * 14 * n + 4 * 8 * ((n + kTpb - 1) / kTpb) + 4096;
  if (precision == GSH_PREC_MIXED) b += 4 * nnz + 8 * n + 4 * 7 * n + 4 * 8 * ((n + kTpb - 1) / kTpb);
  if (precond == GSH_PRECOND_SSOR) {
    // colour-major layout: ELL entries (column, source, value; at most 2.5x the off-diagonal
    // count with padding), 12 vectors and 2 index arrays over rows padded to whole tiles
    const int64_t vt = precision == GSH_PREC_MIXED ? 4 : 8, np = n + 16 * kTpb * 2;
    b += (int64_t)(2.5 * (double)(nnz - n)) * (8 + vt) + np * (12 * vt + 8) + 64 * ((np + kTpb - 1) / kTpb);
  }
  return b;
}

// Retires a request: its solution to the host, its solver back to the mesh's pool.
// A page-locked host vector of mesh mesh_id's length, from the server's pool.
static double* host_vec_get(gsh_server* srv, int32_t mesh_id) {
  auto& f = srv->hfree[mesh_id];
  if (!f.empty()) {
    double* p = f.back();
    f.pop_back();
    return p;
  }
  double* p = nullptr;
  const size_t nb = sizeof(double) * (size_t)srv->meshes[mesh_id]->hm.n;
  const hipError_t e = hipHostMalloc((void**)&p, nb, 0);
  if (e != hipSuccess) throw Error(GSH_E_OOM, "page-locked host vector");
  // the host mapping of the pages faults in at the first host access: take those faults now,
  // not at the copy of a request that reuses a vector the device filled
  std::memset(p, 0, nb);
  return p;
}

static void host_vec_put(gsh_server* srv, int32_t mesh_id, double*& p) {
  if (p) srv->hfree[mesh_id].push_back(p);
  p = nullptr;
}

// The request's solve has ended: its x goes to a page-locked host vector on the solver stream.
// server_finish_retired completes the retirement after one wait for all of a step's copies.
static void server_retire(gsh_server* srv, Request* r, int32_t iters) {
  gsh_solver* sol = r->solver;
  r->hx = host_vec_get(srv, r->mesh_id);
  HIPCHK(hipMemcpyAsync(r->hx, sol->x, sizeof(double) * sol->n, hipMemcpyDeviceToHost, srv->ctx->stream));
  r->iters = iters;
  r->active = false;
  srv->retiring.push_back(r);
}

static void server_finish_retired(gsh_server* srv) {
  if (srv->retiring.empty()) return;
  HIPCHK(hipStreamSynchronize(srv->ctx->stream));
  const auto now = Clock::now();
  for (Request* r : srv->retiring) {
    r->status = 0;
    r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
    srv->finished.push_back(r->ticket);
    // the admission copy of b ran before the set-up the solve waited for
    host_vec_put(srv, r->mesh_id, r->hb);
    srv->pool[r->mesh_id].push_back(r->solver);
    r->solver = nullptr;
    srv->used -= r->bytes;
  }
  srv->retiring.clear();
}

// One step of an FP64 SSOR server: one window of check_every iterations for every active
// request (set up on the staging stream at admission) over a launch table, the window-end
// test, then the finished requests leave with x = (D~ + U)^-1 y.
static void server_step_ssor(gsh_server* srv) {
  hipStream_t st = srv->ctx->stream;
  const auto& cfg = srv->cfg;
  const int na = (int)srv->active.size();
  if (na == 0) return;
  const int rpt = server_rpt(cfg);
  std::vector<gsh_solver*> ss(na);
  std::vector<SsorSys> hsys(na);
  for (int i = 0; i < na; ++i) {
    ss[i] = srv->active[i]->solver;
    hsys[i] = ss[i]->ssor->hs;
  }
  const SsorTable tb = ssor_table(ss.data(), na);
  ensure_cap(srv->dssor, srv->dssor_cap, na);
  ensure_cap(srv->drefs, srv->drefs_cap, (int)tb.refs.size());
  ensure_cap(srv->dgather, srv->dgather_cap, na);
  HIPCHK(hipMemcpyAsync(srv->dssor, hsys.data(), sizeof(SsorSys) * na, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(srv->drefs, tb.refs.data(), sizeof(SsorRef) * tb.refs.size(), hipMemcpyHostToDevice, st));
  for (int i = 0; i < cfg.check_every; ++i)
    ssor_table_sweeps<double>(srv->dssor, srv->drefs, tb, i == cfg.check_every - 1 ? kSsorChk : 0, rpt, st);
  launch_ssor_check_batch(srv->dssor, na, st);
  launch_ssor_gather_states(srv->dssor, na, srv->dgather, st);
  std::vector<PcgState> sts(na);
  HIPCHK(hipMemcpyAsync(sts.data(), srv->dgather, sizeof(PcgState) * na, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  std::vector<Request*> keep;
  for (int i = 0; i < na; ++i) {
    Request* r = srv->active[i];
    r->run += cfg.check_every;
    const PcgState& q = sts[i];
    const bool done = q.done >= 0, maxed = r->run >= cfg.maxiter;
    if (!done && !maxed) {
      keep.push_back(r);
      continue;
    }
    ssor_finish<double>(r->solver, st);
    server_retire(srv, r, done ? q.done : cfg.maxiter);
  }
  srv->active.swap(keep);
}

// One step of a mixed-precision server: FP64 residual checks of the requests due one, then one
// window of FP32 inner PCG iterations for every request with an inner solve running. A request
// follows the same sequence of operations as gsh_solver_solve in mixed precision, so its result
// does not depend on the other requests.
static void server_step_mixed(gsh_server* srv, const std::vector<Request*>& fresh) {
  gsh_ctx* ctx = srv->ctx;
  hipStream_t st = ctx->stream;
  const auto& cfg = srv->cfg;
  for (Request* r : fresh) {
    gsh_solver* s = r->solver;  // values cast and x cleared at admission
    s->hs32.ntiles = s->ntiles;
    r->phase = 0;
    r->outer = 0;
    r->inner_total = 0;
    r->atol2 = -1.0;
  }
  // ---- FP64 residual checks
  std::vector<Request*> chk;
  for (Request* r : srv->active)
    if (r->phase == 0) chk.push_back(r);
  std::vector<Request*> done_now;
  if (!chk.empty()) {
    const int nc = (int)chk.size();
    std::vector<PcgSys<double>> h64(nc);
    std::vector<TileRef> tl;
    for (int j = 0; j < nc; ++j) {
      gsh_solver* s = chk[j]->solver;
      h64[j] = s->hs;
      h64[j].x = s->x;
      h64[j].r = s->r64;
      for (int t = 0; t < s->ntiles; ++t) tl.push_back({j, t});
    }
    ensure_cap(srv->dsys, srv->dsys_cap, nc);
    ensure_cap(srv->dtiles, srv->dtiles_cap, (int)tl.size());
    ensure_cap(srv->dred, srv->dred_cap, 2 * nc);
    HIPCHK(hipMemcpyAsync(srv->dsys, h64.data(), sizeof(PcgSys<double>) * nc, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(srv->dtiles, tl.data(), sizeof(TileRef) * tl.size(), hipMemcpyHostToDevice, st));
    launch_residual_batch(srv->dsys, srv->dtiles, (int)tl.size(), server_rpt(cfg), st);
    launch_reduce_systems(srv->dsys, nc, srv->dred, st);
    std::vector<double> red(2 * nc);
    HIPCHK(hipMemcpyAsync(red.data(), srv->dred, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int j = 0; j < nc; ++j) {
      Request* r = chk[j];
      const double rr = red[2 * j];
      r->outer += 1;
      if (r->atol2 < 0.0) {
        const double atol = cfg.rtol * std::sqrt(red[2 * j + 1]);
        r->atol2 = atol * atol;
        if (atol == 0.0) {  // pcg_solver.py:1089-1095
          launch_fill_f64(r->solver->x, 0.0, r->solver->n, st);
          done_now.push_back(r);
          continue;
        }
      }
      if (rr <= r->atol2 || r->inner_total >= cfg.maxiter || r->outer > 30) {
        done_now.push_back(r);
      } else {
        r->in_rtol = std::max(r->solver->inner_rtol, 0.5 * std::sqrt(r->atol2 / rr));
        r->phase = 1;
        r->inner_fresh = true;
        r->run = 0;
      }
    }
  }
  // ---- one window of inner iterations (fresh inner solves last)
  std::vector<Request*> inr;
  for (Request* r : srv->active)
    if (r->phase == 1 && !r->inner_fresh) inr.push_back(r);
  const int n_old = (int)inr.size();
  for (Request* r : srv->active)
    if (r->phase == 1 && r->inner_fresh) inr.push_back(r);
  if (!inr.empty() && cfg.precond == GSH_PRECOND_SSOR) {
    // FP32 SSOR inner solves: the fresh ones start (their first one gathers the FP32 ELL
    // values), then one window of sweeps for all over a launch table
    const int ni = (int)inr.size();
    std::vector<gsh_solver*> ss(ni);
    std::vector<SsorSysT<float>> h32(ni);
    for (int j = 0; j < ni; ++j) {
      Request* r = inr[j];
      gsh_solver* s = r->solver;
      ss[j] = s;
      if (j >= n_old) {
        ssor_sync_desc32(s, s->r64);
        ssor_begin<float>(s, r->in_rtol, cfg.maxiter, cfg.check_every, !r->ssor_gathered, true, st);
        r->ssor_gathered = true;
      }
      h32[j] = s->ssor->hs32;
    }
    const SsorTable tb = ssor_table(ss.data(), ni);
    const int rpt = server_rpt(cfg);
    ensure_cap(srv->dssor32, srv->dssor32_cap, ni);
    ensure_cap(srv->drefs, srv->drefs_cap, (int)tb.refs.size());
    ensure_cap(srv->dgather, srv->dgather_cap, ni);
    HIPCHK(hipMemcpyAsync(srv->dssor32, h32.data(), sizeof(SsorSysT<float>) * ni, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(srv->drefs, tb.refs.data(), sizeof(SsorRef) * tb.refs.size(), hipMemcpyHostToDevice, st));
    for (int i = 0; i < cfg.check_every; ++i)
      ssor_table_sweeps<float>(srv->dssor32, srv->drefs, tb, i == cfg.check_every - 1 ? kSsorChk : 0, rpt, st);
    launch_ssor_check_batch(srv->dssor32, ni, st);
    launch_ssor_gather_states(srv->dssor32, ni, srv->dgather, st);
    std::vector<PcgState> sts(ni);
    HIPCHK(hipMemcpyAsync(sts.data(), srv->dgather, sizeof(PcgState) * ni, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int j = 0; j < ni; ++j) {
      Request* r = inr[j];
      r->inner_fresh = false;
      r->run += cfg.check_every;
      const bool done = sts[j].done >= 0, maxed = r->run >= cfg.maxiter;
      if (!done && !maxed) continue;
      const int it = done ? sts[j].done : cfg.maxiter;
      ssor_finish<float>(r->solver, st);  // the FP32 correction in the original row order
      launch_refine_update(r->solver->x, r->solver->hs32.x, r->solver->n, st);
      r->inner_total += it;
      r->phase = 0;
      if (it == 0) done_now.push_back(r);
    }
  } else if (!inr.empty()) {
    const int ni = (int)inr.size();
    std::vector<PcgSys<float>> h32(ni);
    std::vector<TileRef> tl, ftl;
    std::vector<double> rts;
    for (int j = 0; j < ni; ++j) {
      gsh_solver* s = inr[j]->solver;
      h32[j] = s->hs32;
      for (int t = 0; t < s->ntiles; ++t) {
        tl.push_back({j, t});
        if (j >= n_old) ftl.push_back({j, t});
      }
      if (j >= n_old) rts.push_back(inr[j]->in_rtol);
    }
    ensure_cap(srv->dsys32, srv->dsys32_cap, ni);
    ensure_cap(srv->dtiles, srv->dtiles_cap, (int)tl.size());
    ensure_cap(srv->dgather, srv->dgather_cap, ni);
    HIPCHK(hipMemcpyAsync(srv->dsys32, h32.data(), sizeof(PcgSys<float>) * ni, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(srv->dtiles, tl.data(), sizeof(TileRef) * tl.size(), hipMemcpyHostToDevice, st));
    if (!ftl.empty()) {
      ensure_cap(srv->dfresh, srv->dfresh_cap, (int)ftl.size());
      ensure_cap(srv->dred, srv->dred_cap, (int)rts.size());
      HIPCHK(hipMemcpyAsync(srv->dfresh, ftl.data(), sizeof(TileRef) * ftl.size(), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(srv->dred, rts.data(), sizeof(double) * rts.size(), hipMemcpyHostToDevice, st));
      launch_pcg_reset<float>(srv->dsys32 + n_old, ni - n_old, cfg.rtol, cfg.maxiter, cfg.check_every, st, srv->dred);
      launch_pcg_init<float>(srv->dsys32, srv->dfresh, (int)ftl.size(), server_rpt(cfg), st);
    }
    for (int i = 0; i < cfg.check_every; ++i)
      launch_pcg_iter<float>(srv->dsys32, srv->dtiles, (int)tl.size(), server_rpt(cfg), st);
    launch_pcg_check<float>(srv->dsys32, ni, st);
    hipLaunchKernelGGL(k_gather_states<float>, dim3((ni + 63) / 64), dim3(64), 0, st, srv->dsys32, ni, srv->dgather);
    std::vector<PcgState> sts(ni);
    HIPCHK(hipMemcpyAsync(sts.data(), srv->dgather, sizeof(PcgState) * ni, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int j = 0; j < ni; ++j) {
      Request* r = inr[j];
      r->inner_fresh = false;
      r->run += cfg.check_every;
      const bool done = sts[j].done >= 0, maxed = r->run >= cfg.maxiter;
      if (!done && !maxed) continue;
      const int it = done ? sts[j].done : cfg.maxiter;
      launch_refine_update(r->solver->x, r->solver->hs32.x, r->solver->n, st);
      r->inner_total += it;
      r->phase = 0;
      if (it == 0) done_now.push_back(r);
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  if (!done_now.empty()) {
    std::vector<Request*> keep;
    for (Request* r : srv->active)
      if (std::find(done_now.begin(), done_now.end(), r) == done_now.end()) keep.push_back(r);
    for (Request* r : done_now) server_retire(srv, r, r->inner_total);
    srv->active.swap(keep);
  }
}


extern "C" {

int gsh_server_create(gsh_ctx* ctx, const gsh_server_config* cfg, gsh_server** out) {
  GSH_API_BODY({
    GSH_REQUIRE(ctx && cfg && out, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(cfg->precision == GSH_PREC_FP64 || cfg->precision == GSH_PREC_MIXED, GSH_E_ARG,
                "unknown precision");
    GSH_REQUIRE(cfg->precond == GSH_PRECOND_JACOBI || cfg->precond == GSH_PRECOND_SSOR, GSH_E_ARG,
                "unknown preconditioner");
    GSH_REQUIRE(cfg->check_every >= 1 && cfg->maxiter >= 1 && cfg->maxiter % cfg->check_every == 0, GSH_E_ARG,
                "maxiter must be a positive multiple of check_every");
    GSH_REQUIRE(cfg->rtol >= 0.0, GSH_E_ARG, "rtol must be >= 0");
    ctx->set();
    auto s = std::make_unique<gsh_server>();
    s->ctx = ctx;
    s->cfg = *cfg;
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    s->budget = cfg->hbm_budget_bytes > 0 ? cfg->hbm_budget_bytes : (int64_t)f;
    HIPCHK(hipStreamCreateWithFlags(&s->stage, hipStreamNonBlocking));
    ctx_retain(ctx);
    *out = s.release();
  })
}

int gsh_server_destroy(gsh_server* srv) {
  if (!srv) return GSH_OK;
  gsh_ctx* ctx = srv->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (srv->stage) (void)hipStreamSynchronize(srv->stage);
  delete srv;
  ctx_release(ctx);
  return GSH_OK;
}

int gsh_server_add_mesh(gsh_server* srv, gsh_mesh* mesh, int32_t* id) {
  GSH_API_BODY({
    GSH_REQUIRE(srv && mesh && id, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(mesh->ctx == srv->ctx, GSH_E_ARG, "mesh belongs to another context");
    gsh_step_params p;
    gsh_step_params_default(&p);
    srv->meshes.push_back(mesh);
    srv->params.push_back(p);
    srv->pool.emplace_back();
    srv->hfree.emplace_back();
    *id = (int32_t)srv->meshes.size() - 1;
  })
}

int gsh_server_set_mesh_params(gsh_server* srv, int32_t mesh_id, const gsh_step_params* p) {
  GSH_API_BODY({
    GSH_REQUIRE(srv && p, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(mesh_id >= 0 && mesh_id < (int32_t)srv->meshes.size(), GSH_E_ARG, "unknown mesh id");
    (void)step_args(p);  // validated now, not at admission
    srv->params[mesh_id] = *p;
  })
}

int gsh_server_submit(gsh_server* srv, int32_t mesh_id, const double* b, int with_bt, int64_t* ticket) {
  GSH_API_BODY({
    GSH_REQUIRE(srv && ticket, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(mesh_id >= 0 && mesh_id < (int32_t)srv->meshes.size(), GSH_E_ARG, "unknown mesh id");
    auto r = std::make_unique<Request>();
    r->ticket = srv->next_ticket++;
    r->mesh_id = mesh_id;
    r->n = srv->meshes[mesh_id]->hm.n;
    r->with_bt = with_bt != 0;
    if (b) {
      r->has_b = true;
      r->hb = host_vec_get(srv, mesh_id);
      std::memcpy(r->hb, b, sizeof(double) * (size_t)r->n);
    }
    r->t_submit = Clock::now();
    r->bytes = solver_bytes(srv->meshes[mesh_id], srv->cfg.precision, srv->cfg.precond);
    *ticket = r->ticket;
    srv->queue.push_back(r.get());
    srv->reqs[r->ticket] = std::move(r);
  })
}

int gsh_server_step(gsh_server* srv, int32_t* active_out, int32_t* queued_out) {
  GSH_API_BODY({
    GSH_REQUIRE(srv, GSH_E_ARG, "server is NULL");
    gsh_ctx* ctx = srv->ctx;
    ctx->set();
    hipStream_t st = ctx->stream;
    const auto& cfg = srv->cfg;
    // ---- requests whose set-up has finished join now (the oldest one is waited for when
    // nothing else could run)
    std::vector<Request*> fresh;
    {
      std::vector<Request*> still;
      for (Request* r : srv->staged) {
        const hipError_t q = hipEventQuery(r->ready);
        if (q == hipSuccess || (q == hipErrorNotReady && srv->active.empty() && fresh.empty())) {
          if (q != hipSuccess) HIPCHK(hipStreamWaitEvent(st, r->ready, 0));
          fresh.push_back(r);
        } else if (q == hipErrorNotReady) {
          still.push_back(r);
        } else {
          hip_check(q, "request set-up");
        }
      }
      srv->staged.swap(still);
    }
    // ---- admission; the set-up goes to the staging stream
    hipStream_t sg = srv->stage;
    while (!srv->queue.empty()) {
      Request* r = srv->queue.front();
      const int in_flight = (int)(srv->active.size() + fresh.size() + srv->staged.size());
      if (cfg.max_active > 0 && in_flight >= cfg.max_active) break;
      if (srv->used + r->bytes > srv->budget) {
        if (in_flight == 0) {  // can never fit
          r->status = GSH_E_ADMISSION;
          host_vec_put(srv, r->mesh_id, r->hb);
          srv->queue.pop_front();
          continue;
        }
        break;
      }
      srv->queue.pop_front();
      gsh_mesh* m = srv->meshes[r->mesh_id];
      auto& idle = srv->pool[r->mesh_id];
      gsh_solver* s = nullptr;
      if (!idle.empty()) {
        s = idle.back();
        idle.pop_back();
      } else {
        try {
          s = mesh_solver_new(m, cfg.precision, cfg.precond, server_rpt(cfg));
        } catch (const Error& e) {
          throw Error(GSH_E_OOM, std::string("server admission: ") + e.what());
        }
      }
      for (auto* hh : s->h) launch_fill_f64(hh, 0.0, s->n, sg);  // each request is a cold start
      // the mesh's step parameters (gsh_server_set_mesh_params; default gsh_step_params_default)
      const StepArgs a = step_args(&srv->params[r->mesh_id]);
      if (r->with_bt) {
        launch_backtrack(m->d, a, nullptr, nullptr, 0, sg);
        launch_rhs(m->d, a, sg);
      }
      // the request's matrix (and right-hand side unless given) from the mesh state
      launch_assemble(m->d, a, s->val, s->b, sg);
      if (r->has_b) HIPCHK(hipMemcpyAsync(s->b, r->hb, sizeof(double) * s->n, hipMemcpyHostToDevice, sg));
      if (cfg.precision == GSH_PREC_MIXED) {
        launch_cast_values(s->val, s->val32, s->nnz, sg);
        launch_fill_f64(s->x, 0.0, s->n, sg);
      } else if (cfg.precond == GSH_PRECOND_SSOR) {
        // the whole SSOR set-up (values into the colour-major layout, r0, the set-up sweeps)
        ssor_sync_desc(s, sg);
        ssor_begin<double>(s, cfg.rtol, cfg.maxiter, cfg.check_every, true, true, sg);
      }
      s->rpt = server_rpt(cfg);
      s->ntiles = s->tiles_for(s->n);
      s->hs.x = s->x;
      s->hs.h0 = s->h[0];
      s->hs.h1 = s->h[1];
      s->hs.h2 = s->h[2];
      s->hs.ntiles = s->ntiles;
      r->solver = s;
      r->active = true;
      r->run = 0;
      srv->used += r->bytes;
      if (!r->ready) HIPCHK(hipEventCreateWithFlags(&r->ready, hipEventDisableTiming));
      HIPCHK(hipEventRecord(r->ready, sg));
      srv->staged.push_back(r);
    }
    for (auto* r : fresh) srv->active.push_back(r);
    const bool jacobi64 = cfg.precision == GSH_PREC_FP64 && cfg.precond == GSH_PRECOND_JACOBI;
    const int na = jacobi64 ? (int)srv->active.size() : 0;
    if (cfg.precision == GSH_PREC_MIXED) {
      server_step_mixed(srv, fresh);
    } else if (cfg.precond == GSH_PRECOND_SSOR) {
      server_step_ssor(srv);
    }
    if (na > 0) {
      // ---- descriptors and tile tables (fresh systems are the last ones)
      std::vector<PcgSys<double>> hs(na);
      std::vector<TileRef> tiles, ftiles;
      for (int i = 0; i < na; ++i) {
        gsh_solver* s = srv->active[i]->solver;
        hs[i] = s->hs;
        const bool is_fresh = i >= na - (int)fresh.size();
        for (int t = 0; t < s->ntiles; ++t) {
          tiles.push_back({i, t});
          if (is_fresh) ftiles.push_back({i, t});
        }
      }
      ensure_cap(srv->dsys, srv->dsys_cap, na);
      ensure_cap(srv->dtiles, srv->dtiles_cap, (int)tiles.size());
      ensure_cap(srv->dgather, srv->dgather_cap, na);
      HIPCHK(hipMemcpyAsync(srv->dsys, hs.data(), sizeof(PcgSys<double>) * na, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(srv->dtiles, tiles.data(), sizeof(TileRef) * tiles.size(), hipMemcpyHostToDevice, st));
      if (!fresh.empty()) {
        ensure_cap(srv->dfresh, srv->dfresh_cap, (int)ftiles.size());
        HIPCHK(hipMemcpyAsync(srv->dfresh, ftiles.data(), sizeof(TileRef) * ftiles.size(), hipMemcpyHostToDevice,
                              st));
        const int f0 = na - (int)fresh.size();
        launch_pcg_reset<double>(srv->dsys + f0, (int)fresh.size(), cfg.rtol, cfg.maxiter, cfg.check_every, st);
        launch_pcg_init<double>(srv->dsys, srv->dfresh, (int)ftiles.size(), server_rpt(cfg), st);
      }
      // ---- one check window for every active system
      for (int i = 0; i < cfg.check_every; ++i) launch_pcg_iter<double>(srv->dsys, srv->dtiles, (int)tiles.size(), server_rpt(cfg), st);
      launch_pcg_check<double>(srv->dsys, na, st);
      for (auto* r : srv->active) r->run += cfg.check_every;
      hipLaunchKernelGGL(k_gather_states<double>, dim3((na + 63) / 64), dim3(64), 0, st, srv->dsys, na, srv->dgather);
      std::vector<PcgState> sts(na);
      HIPCHK(hipMemcpyAsync(sts.data(), srv->dgather, sizeof(PcgState) * na, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipGetLastError());
      // ---- retire finished systems
      std::vector<Request*> keep;
      for (int i = 0; i < na; ++i) {
        Request* r = srv->active[i];
        const PcgState& s = sts[i];
        const bool done = s.done >= 0;
        const bool maxed = r->run >= cfg.maxiter;
        if (!done && !maxed) {
          keep.push_back(r);
          continue;
        }
        if (s.zero) launch_fill_f64(r->solver->x, 0.0, r->solver->n, st);
        server_retire(srv, r, done ? s.done : cfg.maxiter);
      }
      srv->active.swap(keep);
    }
    server_finish_retired(srv);
    if (active_out) *active_out = (int32_t)(srv->active.size() + srv->staged.size());
    if (queued_out) *queued_out = (int32_t)srv->queue.size();
  })
}

int gsh_server_finished(gsh_server* srv, int64_t* tickets, int32_t cap, int32_t* count) {
  GSH_API_BODY({
    GSH_REQUIRE(srv && tickets && count && cap >= 0, GSH_E_ARG, "bad arguments");
    int32_t k = 0;
    while (k < cap && !srv->finished.empty()) {
      tickets[k++] = srv->finished.front();
      srv->finished.pop_front();
    }
    *count = k;
  })
}

int gsh_server_release(gsh_server* srv, int64_t ticket) {
  GSH_API_BODY({
    GSH_REQUIRE(srv, GSH_E_ARG, "server is NULL");
    auto it = srv->reqs.find(ticket);
    GSH_REQUIRE(it != srv->reqs.end(), GSH_E_NOTFOUND, "unknown ticket");
    Request* r = it->second.get();
    GSH_REQUIRE(r->status != 1, GSH_E_STATE, "request is still queued or running");
    host_vec_put(srv, r->mesh_id, r->hx);
    host_vec_put(srv, r->mesh_id, r->hb);
    srv->reqs.erase(it);
  })
}

int gsh_server_poll(gsh_server* srv, int64_t ticket, double* x_host, int32_t* iters_out, int32_t* status_out,
                    double* latency_ms_out) {
  GSH_API_BODY({
    GSH_REQUIRE(srv, GSH_E_ARG, "server is NULL");
    auto it = srv->reqs.find(ticket);
    GSH_REQUIRE(it != srv->reqs.end(), GSH_E_NOTFOUND, "unknown ticket");
    Request* r = it->second.get();
    if (status_out) *status_out = r->status;
    if (r->status == 0) {
      if (x_host) std::memcpy(x_host, r->hx, sizeof(double) * (size_t)r->n);
      if (iters_out) *iters_out = r->iters;
      if (latency_ms_out) *latency_ms_out = r->latency_ms;
    }
  })
}

}  // extern "C"

// ====================================================================== multi-RHS block
// Ensemble members of one mesh solved together on its shared CSR pattern (member-interleaved
// layout, block_kernels.hip).
struct gsh_block {
  gsh_ctx* ctx = nullptr;
  gsh_mesh* mesh = nullptr;
  int32_t members = 0, n = 0;
  int64_t nnz = 0;
  DevPool pool;
  double *tmp_val = nullptr, *tmp_b = nullptr;  // one member's assembled values and RHS
  double *vblk = nullptr, *bblk = nullptr;      // [nnz][M], [n][M]
  BlkSys<double> S{};
  std::vector<BlkState> last;                   // member states of the last solve
};

static int32_t blk_width(int32_t members) {
  int32_t M = 1;
  while (M < members) M *= 2;
  return M;
}

extern "C" {

int gsh_block_create(gsh_mesh* mesh, int32_t members, gsh_block** out) {
  GSH_API_BODY({
    GSH_REQUIRE(mesh && out, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(mesh->ctx, GSH_E_STATE, "host-only mesh (created without a context)");
    GSH_REQUIRE(members >= 1 && members <= 256, GSH_E_ARG, "members must be in [1, 256]");
    mesh->ctx->set();
    auto b = std::make_unique<gsh_block>();
    b->ctx = mesh->ctx;
    b->mesh = mesh;
    b->members = members;
    const Topology& t = mesh->topo;
    const int32_t n = mesh->hm.n, M = blk_width(members);
    const int64_t nnz = (int64_t)t.indices.size();
    b->n = n;
    b->nnz = nnz;
    DevPool& P = b->pool;
    BlkSys<double>& S = b->S;
    S.n = n;
    S.M = M;
    S.ntiles = (n + kBlkTile - 1) / kBlkTile;
    S.rowptr = P.upload(t.indptr);
    S.col = P.upload(t.indices);
    S.diag = P.upload(diag_abs_of(t.indptr, t.indices, n));
    b->vblk = P.alloc<double>(nnz * M);
    b->bblk = P.alloc<double>((int64_t)n * M);
    S.val = b->vblk;
    S.b = b->bblk;
    for (double** v : {&S.x, &S.r, &S.z, &S.pa, &S.pb, &S.q, &S.invd}) *v = P.alloc<double>((int64_t)n * M);
    S.part_a = P.alloc<double>((int64_t)S.ntiles * M);
    S.part_b = P.alloc<double>((int64_t)S.ntiles * M);
    S.st = P.alloc<BlkState>(M);
    b->tmp_val = P.alloc<double>(nnz);
    b->tmp_b = P.alloc<double>(n);
    // members beyond `members` (the block is a power of two wide): unit values, zero RHS, so they
    // stop at once with x = 0
    launch_fill_f64(b->vblk, 1.0, nnz * M, b->ctx->stream);
    HIPCHK(hipStreamSynchronize(b->ctx->stream));
    ctx_retain(b->ctx);
    *out = b.release();
  })
}

int gsh_block_destroy(gsh_block* b) {
  if (!b) return GSH_OK;
  gsh_ctx* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete b;
  ctx_release(ctx);
  return GSH_OK;
}

int gsh_block_assemble(gsh_block* b, int32_t member, const gsh_step_params* p, double time) {
  (void)time;  // as gsh_assemble
  GSH_API_BODY({
    GSH_REQUIRE(b && p, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(member >= 0 && member < b->members, GSH_E_ARG, "member out of range");
    b->ctx->set();
    hipStream_t st = b->ctx->stream;
    launch_assemble(b->mesh->d, step_args(p), b->tmp_val, b->tmp_b, st);
    launch_blk_put(b->vblk, b->bblk, b->tmp_val, b->tmp_b, b->nnz, b->n, b->S.M, member, st);
    HIPCHK(hipGetLastError());
  })
}

int gsh_block_set_member(gsh_block* b, int32_t member, const double* values_host, const double* b_host) {
  GSH_API_BODY({
    GSH_REQUIRE(b, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(member >= 0 && member < b->members, GSH_E_ARG, "member out of range");
    b->ctx->set();
    hipStream_t st = b->ctx->stream;
    if (values_host) HIPCHK(hipMemcpyAsync(b->tmp_val, values_host, 8 * b->nnz, hipMemcpyHostToDevice, st));
    if (b_host) HIPCHK(hipMemcpyAsync(b->tmp_b, b_host, 8 * (size_t)b->n, hipMemcpyHostToDevice, st));
    launch_blk_put(b->vblk, b->bblk, values_host ? b->tmp_val : nullptr, b_host ? b->tmp_b : nullptr, b->nnz, b->n,
                   b->S.M, member, st);
    HIPCHK(hipStreamSynchronize(st));
  })
}

int gsh_block_solve(gsh_block* b, double rtol, int maxiter, int check_every, double timeout_s, int32_t* iters_out,
                    int32_t* status_out) {
  GSH_API_BODY({
    GSH_REQUIRE(b, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(rtol >= 0.0, GSH_E_ARG, "rtol must be >= 0");
    if (maxiter <= 0) maxiter = 10 * b->n;  // pcg_solver.py:1082-1083
    if (check_every < 1) check_every = 1;
    if (timeout_s <= 0.0) timeout_s = 60.0;
    gsh_ctx* ctx = b->ctx;
    ctx->set();
    hipStream_t st = ctx->stream;
    const BlkSys<double>& S = b->S;
    const int M = S.M;
    b->last.assign(M, BlkState{});
    // one readback slot of M states per check window
    BlkState* host = reinterpret_cast<BlkState*>(ctx->pinned_slots((M * (int)sizeof(BlkState) + sizeof(PcgState) - 1) /
                                                                   sizeof(PcgState)));
    launch_blk_init(S, st);
    launch_blk_reduce_init(S, rtol, st);
    HIPCHK(hipGetLastError());
    const auto t0 = Clock::now();
    int it = 0;
    bool all_done = false;
    while (it < maxiter && !all_done) {
      const int w = std::min(check_every, maxiter - it);
      for (int i = 0; i < w; ++i, ++it) launch_blk_iter(S, it, rtol, maxiter, check_every, st);
      HIPCHK(hipMemcpyAsync(host, S.st, sizeof(BlkState) * M, hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(ctx->ev[0], st));
      wait_event(ctx->ev[0], timeout_s, "block PCG window");
      all_done = true;
      for (int m = 0; m < b->members; ++m) all_done = all_done && host[m].done >= 0;
      if (std::chrono::duration<double>(Clock::now() - t0).count() > timeout_s)
        throw Error(GSH_E_TIMEOUT, "block PCG: time limit exceeded");
      HIPCHK(hipGetLastError());
    }
    b->last.assign(host, host + M);
    for (int m = 0; m < b->members; ++m) {
      const BlkState& s = b->last[m];
      if (iters_out) iters_out[m] = s.done >= 0 ? s.done : maxiter;
      if (status_out) status_out[m] = s.done >= 0 ? 0 : GSH_W_MAXITER;
    }
  })
}

int gsh_block_get_solution(gsh_block* b, int32_t member, double* x_host) {
  GSH_API_BODY({
    GSH_REQUIRE(b && x_host, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(member >= 0 && member < b->members, GSH_E_ARG, "member out of range");
    b->ctx->set();
    hipStream_t st = b->ctx->stream;
    launch_blk_get(b->tmp_b, b->S.x, b->n, b->S.M, member, st);
    HIPCHK(hipMemcpyAsync(x_host, b->tmp_b, 8 * (size_t)b->n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  })
}

}  // extern "C"

// ====================================================================== measurement helpers
namespace gsh {
__global__ void k_copy_probe(const double4* __restrict__ src, double4* __restrict__ dst, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

// U records per thread per round (U loads in flight before the stores); NT: stores bypass the
// caches (nontemporal). Every index is below n4.
template <int U, bool NT>
__global__ void k_copy_probe_u(const double4* __restrict__ src, double4* __restrict__ dst, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n4; i += U * stride) {
    double4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NT) {
        __builtin_nontemporal_store(v[u].x, &dst[i + u * stride].x);
        __builtin_nontemporal_store(v[u].y, &dst[i + u * stride].y);
        __builtin_nontemporal_store(v[u].z, &dst[i + u * stride].z);
        __builtin_nontemporal_store(v[u].w, &dst[i + u * stride].w);
      } else {
        dst[i + u * stride] = v[u];
      }
    }
  }
  for (; i < n4; i += stride) dst[i] = src[i];
}
}  // namespace gsh

extern "C" {

int gsh_copy_probe(gsh_ctx* ctx, int64_t bytes, int reps, double* ms_avg) {
  GSH_API_BODY({
    GSH_REQUIRE(ctx && ms_avg && bytes >= 32 && reps >= 1, GSH_E_ARG, "bad arguments");
    ctx->set();
    hipStream_t st = ctx->stream;
    const int64_t n4 = bytes / 32;
    double4 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n4 * 32));
    HIPCHK(hipMalloc(&b, n4 * 32));
    HIPCHK(hipMemsetAsync(a, 0, n4 * 32, st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int dev_cus = 0;
    HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int cus = std::max(1, dev_cus);
    // the fastest of a few launch shapes: the achievable-bandwidth reference
    double best = 1e30;
    for (int variant = 0; variant < 7; ++variant) {
      const int mult[7] = {16, 8, 16, 32, 8, 16, 32};
      const int grid = cus * mult[variant];
      auto launch = [&]() {
        switch (variant) {
          case 0: hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, st, a, b, n4); break;
          case 1:
          case 2:
          case 3: hipLaunchKernelGGL((k_copy_probe_u<4, false>), dim3(grid), dim3(256), 0, st, a, b, n4); break;
          default: hipLaunchKernelGGL((k_copy_probe_u<4, true>), dim3(grid), dim3(256), 0, st, a, b, n4); break;
        }
      };
      launch();  // warm-up
      HIPCHK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; ++r) launch();
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, (double)ms / reps);
    }
    *ms_avg = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
  })
}

int gsh_solver_profile(gsh_solver* s, double rtol, int maxiter, int check_every, double* ms, int* iters_out) {
  GSH_API_BODY({
    GSH_REQUIRE(s && ms, GSH_E_ARG, "NULL argument");
    GSH_REQUIRE(s->precision == GSH_PREC_FP64, GSH_E_ARG, "phase profiling runs the FP64 solver");
    if (maxiter <= 0) maxiter = 10 * s->n;
    s->ctx->set();
    for (int i = 0; i < 4; ++i) ms[i] = 0.0;
    if (s->precond == GSH_PRECOND_SSOR) {
      SsorTimes tt;
      ssor_sync_desc(s);
      const PcgState r = run_ssor<double>(s, rtol, maxiter, check_every, 60.0, &tt);
      HIPCHK(hipStreamSynchronize(s->ctx->stream));
      solver_roll(s);
      if (tt.iters) {
        ms[0] = tt.ms_bwd / tt.iters;
        ms[1] = tt.ms_fwd / tt.iters;
        ms[2] = tt.ms_upd / tt.iters;
      }
      if (iters_out) *iters_out = iters_of(r, maxiter);
      return GSH_OK;
    }
    solver_sync_desc(s);
    PcgOutcome o;
    PhaseTimes pt;
    run_pcg<double>(s->ctx, s->ds, 1, s->dtiles, s->ntiles, s->rpt, rtol, maxiter, check_every, 60.0, o, nullptr,
                    -1, &pt);
    const int it = iters_of(o.st[0], maxiter);
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    solver_roll(s);
    ms[0] = pt.launches ? pt.ms_a / pt.launches : 0.0;
    ms[1] = pt.launches ? pt.ms_b / pt.launches : 0.0;
    if (iters_out) *iters_out = it;
  })
}

}  // extern "C"

extern "C" int gsh_solver_get_rhs(gsh_solver* s, double* b_host) {
  GSH_API_BODY({
    GSH_REQUIRE(s && b_host, GSH_E_ARG, "NULL argument");
    s->ctx->set();
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    HIPCHK(hipMemcpy(b_host, s->b, sizeof(double) * s->n, hipMemcpyDeviceToHost));
  })
}

extern "C" int gsh_solver_set_rhs(gsh_solver* s, const double* b_host) {
  GSH_API_BODY({
    GSH_REQUIRE(s && b_host, GSH_E_ARG, "NULL argument");
    s->ctx->set();
    HIPCHK(hipMemcpyAsync(s->b, b_host, sizeof(double) * s->n, hipMemcpyHostToDevice, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
  })
}

extern "C" int gsh_solver_get_solution(gsh_solver* s, double* x_host) {
  GSH_API_BODY({
    GSH_REQUIRE(s && x_host, GSH_E_ARG, "NULL argument");
    s->ctx->set();
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    // after a solve the solution is the newest history entry (pointer rotation of rollHistory)
    HIPCHK(hipMemcpy(x_host, s->h[0], sizeof(double) * s->n, hipMemcpyDeviceToHost));
  })
}
