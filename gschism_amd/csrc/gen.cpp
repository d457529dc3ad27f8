// gen.cpp -- synthetic inputs: RNG, jittered-grid Delaunay channel, log-normal bathymetry,
// divergence-free velocity fields. These build the workloads the hot path is measured on;
// the reference reads the same kinds of arrays from hgrid.gr3 / vgrid.in / hotstart files
// (read_and_process_grid.cpp:829-1249).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "abi_util.h"

namespace gsh {

// ------------------------------------------------------------------ RNG
static uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

Rng::Rng(uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
}
uint64_t Rng::next() {
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}
double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
double Rng::uniform_open() { return ((double)(next() >> 11) + 1.0) * 0x1.0p-53; }
double Rng::normal() {
  if (has_spare) {
    has_spare = false;
    return spare;
  }
  const double u1 = uniform_open(), u2 = uniform();
  const double r = std::sqrt(-2.0 * std::log(u1));
  const double a = 2.0 * M_PI * u2;
  spare = r * std::sin(a);
  has_spare = true;
  return r * std::cos(a);
}

// ------------------------------------------------------------------ Delaunay channel
struct Channel {
  int32_t nx = 0, ny = 0;
  std::vector<double> x, y;
  std::vector<int32_t> tri;    // [nt][3] CCW
  std::vector<int8_t> open_bnd;
};

// In-circle determinant of d with respect to CCW triangle (a, b, c); > 0 means inside.
static double incircle(double ax, double ay, double bx, double by, double cx, double cy,
                       double dx, double dy) {
  const double adx = ax - dx, ady = ay - dy;
  const double bdx = bx - dx, bdy = by - dy;
  const double cdx = cx - dx, cdy = cy - dy;
  const double ad = adx * adx + ady * ady;
  const double bd = bdx * bdx + bdy * bdy;
  const double cd = cdx * cdx + cdy * cdy;
  return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
}

static void build_neighbors(int32_t nt, const std::vector<int32_t>& tri, std::vector<int32_t>& nb) {
  // half-edge keys (min,max) -> match pairs by sorting
  std::vector<uint64_t> key((size_t)nt * 3);
  std::vector<int32_t> idx((size_t)nt * 3);
  for (int32_t t = 0; t < nt; ++t)
    for (int e = 0; e < 3; ++e) {
      uint32_t a = (uint32_t)tri[3 * t + (e + 1) % 3], b = (uint32_t)tri[3 * t + (e + 2) % 3];
      if (a > b) std::swap(a, b);
      key[3 * (size_t)t + e] = ((uint64_t)a << 32) | b;
      idx[3 * (size_t)t + e] = 3 * t + e;
    }
  std::vector<int32_t> order((size_t)nt * 3);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int32_t p, int32_t q) {
    return key[p] < key[q] || (key[p] == key[q] && p < q);
  });
  nb.assign((size_t)nt * 3, -1);
  for (size_t i = 0; i + 1 < order.size(); ++i) {
    if (key[order[i]] == key[order[i + 1]]) {
      const int32_t h0 = order[i], h1 = order[i + 1];
      nb[h0] = h1 / 3;
      nb[h1] = h0 / 3;
      ++i;
    }
  }
}

static void lawson_flip(Channel& c, double tol) {
  const int32_t nt = (int32_t)(c.tri.size() / 3);
  std::vector<int32_t> nb;
  build_neighbors(nt, c.tri, nb);
  std::vector<int64_t> stack;
  stack.reserve((size_t)nt * 3);
  for (int32_t t = 0; t < nt; ++t)
    for (int e = 0; e < 3; ++e)
      if (nb[3 * t + e] > t) stack.push_back(3 * (int64_t)t + e);
  auto& T = c.tri;
  auto& X = c.x;
  auto& Y = c.y;
  int64_t guard = 0;
  while (!stack.empty()) {
    const int64_t he = stack.back();
    stack.pop_back();
    const int32_t t = (int32_t)(he / 3);
    const int e = (int)(he % 3);
    const int32_t u = nb[3 * t + e];
    if (u < 0) continue;
    const int32_t a = T[3 * t + e], b = T[3 * t + (e + 1) % 3], cc = T[3 * t + (e + 2) % 3];
    int f = -1;
    for (int k = 0; k < 3; ++k)
      if (T[3 * u + k] != b && T[3 * u + k] != cc) f = k;
    if (f < 0) continue;
    const int32_t d = T[3 * u + f];
    if (T[3 * u + (f + 1) % 3] != cc || T[3 * u + (f + 2) % 3] != b) continue;  // inconsistent
    if (incircle(X[a], Y[a], X[b], Y[b], X[cc], Y[cc], X[d], Y[d]) <= tol) continue;
    if (++guard > (int64_t)nt * 64) throw Error(GSH_E_TOPOLOGY, "Lawson flipping did not converge");
    const int32_t t_ca = nb[3 * t + (e + 1) % 3], t_ab = nb[3 * t + (e + 2) % 3];
    const int32_t u_bd = nb[3 * u + (f + 1) % 3], u_dc = nb[3 * u + (f + 2) % 3];
    // t' = (a, b, d), u' = (a, d, c)
    T[3 * t + 0] = a; T[3 * t + 1] = b; T[3 * t + 2] = d;
    nb[3 * t + 0] = u_bd; nb[3 * t + 1] = u; nb[3 * t + 2] = t_ab;
    T[3 * u + 0] = a; T[3 * u + 1] = d; T[3 * u + 2] = cc;
    nb[3 * u + 0] = u_dc; nb[3 * u + 1] = t_ca; nb[3 * u + 2] = t;
    if (u_bd >= 0)
      for (int k = 0; k < 3; ++k)
        if (nb[3 * u_bd + k] == u) nb[3 * u_bd + k] = t;
    if (t_ca >= 0)
      for (int k = 0; k < 3; ++k)
        if (nb[3 * t_ca + k] == t) nb[3 * t_ca + k] = u;
    stack.push_back(3 * (int64_t)t + 0);
    stack.push_back(3 * (int64_t)t + 2);
    stack.push_back(3 * (int64_t)u + 0);
    stack.push_back(3 * (int64_t)u + 1);
  }
}

static Channel make_channel(int32_t nx, int32_t ny, double dx, double jitter, uint64_t seed) {
  if (nx < 2 || ny < 2) throw Error(GSH_E_ARG, "channel needs nx, ny >= 2");
  if ((int64_t)nx * ny > (int64_t)1 << 30) throw Error(GSH_E_ARG, "channel too large");
  if (!(jitter >= 0.0 && jitter < 0.35)) throw Error(GSH_E_ARG, "jitter must be in [0, 0.35)");
  Channel c;
  c.nx = nx;
  c.ny = ny;
  const int64_t n = (int64_t)nx * ny;
  c.x.resize(n);
  c.y.resize(n);
  c.open_bnd.assign(n, 0);
  Rng rng(seed);
  for (int32_t j = 0; j < ny; ++j)
    for (int32_t i = 0; i < nx; ++i) {
      const int64_t id = (int64_t)j * nx + i;
      // two draws per node in node order, always taken so that the stream stays aligned
      double jx = (2.0 * rng.uniform() - 1.0) * jitter * dx;
      double jy = (2.0 * rng.uniform() - 1.0) * jitter * dx;
      const bool west = (i == 0), east = (i == nx - 1), south = (j == 0), north = (j == ny - 1);
      if (west || east) jx = 0.0;
      if (south || north) jy = 0.0;
      c.x[id] = i * dx + jx;
      c.y[id] = j * dx + jy;
      if (west) c.open_bnd[id] = 1;
    }
  const int64_t nt = 2 * (int64_t)(nx - 1) * (ny - 1);
  c.tri.resize(3 * nt);
  for (int32_t j = 0; j < ny - 1; ++j)
    for (int32_t i = 0; i < nx - 1; ++i) {
      const int32_t a = j * nx + i, b = a + 1, cc = a + nx + 1, d = a + nx;
      const int64_t t = 2 * ((int64_t)j * (nx - 1) + i);
      c.tri[3 * t + 0] = a; c.tri[3 * t + 1] = b; c.tri[3 * t + 2] = cc;
      c.tri[3 * t + 3] = a; c.tri[3 * t + 4] = cc; c.tri[3 * t + 5] = d;
    }
  lawson_flip(c, 1e-9 * dx * dx * dx * dx);
  return c;
}

// ------------------------------------------------------------------ Gaussian random field
static void gaussian_blur_1d(const double* in, double* out, int32_t len, int64_t stride,
                             const std::vector<double>& w, int32_t r) {
  for (int32_t i = 0; i < len; ++i) {
    double s = 0.0;
    for (int32_t k = -r; k <= r; ++k) {
      int32_t j = i + k;
      if (j < 0) j = -j - 1;                // reflect
      if (j >= len) j = 2 * len - j - 1;
      if (j < 0) j = 0;
      if (j >= len) j = len - 1;
      s += w[k + r] * in[(int64_t)j * stride];
    }
    out[(int64_t)i * stride] = s;
  }
}

static void bathymetry(int32_t nx, int32_t ny, double corr, double hmin, double hmax, uint64_t seed,
                       double* depth) {
  if (nx < 1 || ny < 1 || !(corr > 0.0) || !(hmin > 0.0) || !(hmax > hmin))
    throw Error(GSH_E_ARG, "bathymetry: bad arguments");
  const int64_t n = (int64_t)nx * ny;
  std::vector<double> a(n), b(n);
  Rng rng(seed);
  for (int64_t i = 0; i < n; ++i) a[i] = rng.normal();
  // white noise blurred with a Gaussian of std s has covariance exp(-r^2 / (4 s^2));
  // s = corr / sqrt(2) gives the Gaussian covariance exp(-r^2 / (2 corr^2)).
  const double s = corr / std::sqrt(2.0);
  const int32_t r = std::max(1, (int32_t)std::ceil(3.0 * s));
  std::vector<double> w(2 * r + 1);
  double wsum = 0.0;
  for (int32_t k = -r; k <= r; ++k) wsum += (w[k + r] = std::exp(-0.5 * (k * k) / (s * s)));
  for (auto& v : w) v /= wsum;
  for (int32_t j = 0; j < ny; ++j) gaussian_blur_1d(&a[(int64_t)j * nx], &b[(int64_t)j * nx], nx, 1, w, r);
  for (int32_t i = 0; i < nx; ++i) gaussian_blur_1d(&b[i], &a[i], ny, nx, w, r);
  double gmin = a[0], gmax = a[0];
  for (int64_t i = 0; i < n; ++i) gmin = std::min(gmin, a[i]), gmax = std::max(gmax, a[i]);
  const double l0 = std::log(hmin), l1 = std::log(hmax);
  const double span = (gmax > gmin) ? (gmax - gmin) : 1.0;
  for (int64_t i = 0; i < n; ++i) depth[i] = std::exp(l0 + (l1 - l0) * (a[i] - gmin) / span);
}

// ------------------------------------------------------------------ velocity field
struct Modes {
  std::vector<double> kx, ky, ph, amp;
  double scale = 1.0;
};

static Modes make_modes(double lx, double ly, double dx, int32_t nmodes, double umax, uint64_t seed) {
  if (nmodes < 1 || !(dx > 0.0) || !(lx > 0.0) || !(ly > 0.0)) throw Error(GSH_E_ARG, "velocity: bad arguments");
  Modes m;
  Rng rng(seed);
  for (int32_t q = 0; q < nmodes; ++q) {
    const double lam = dx * std::exp(std::log(5.0) + (std::log(50.0) - std::log(5.0)) * rng.uniform());
    const double th = 2.0 * M_PI * rng.uniform();
    const double k = 2.0 * M_PI / lam;
    m.kx.push_back(k * std::cos(th));
    m.ky.push_back(k * std::sin(th));
    m.ph.push_back(2.0 * M_PI * rng.uniform());
    m.amp.push_back(rng.normal() / k);  // mode velocity amplitude ~ N(0, 1)
  }
  // scale so that max |u| over a 512 x 512 sampling of the domain equals umax
  double vmax = 0.0;
  for (int j = 0; j <= 512; ++j)
    for (int i = 0; i <= 512; ++i) {
      const double px = lx * i / 512.0, py = ly * j / 512.0;
      double u = 0.0, v = 0.0;
      for (int32_t q = 0; q < nmodes; ++q) {
        const double c = std::cos(m.kx[q] * px + m.ky[q] * py + m.ph[q]);
        u += m.amp[q] * m.ky[q] * c;
        v -= m.amp[q] * m.kx[q] * c;
      }
      vmax = std::max(vmax, std::sqrt(u * u + v * v));
    }
  m.scale = (vmax > 0.0) ? umax / vmax : 0.0;
  return m;
}

}  // namespace gsh

using namespace gsh;

struct gsh_gen {
  Channel c;
};

extern "C" {

int gsh_gen_channel(int32_t nx, int32_t ny, double dx, double jitter, uint64_t seed, gsh_gen** out) {
  GSH_API_BODY({
    GSH_REQUIRE(out, GSH_E_ARG, "out is NULL");
    GSH_REQUIRE(dx > 0.0, GSH_E_ARG, "dx must be > 0");
    auto* g = new gsh_gen;
    try {
      g->c = make_channel(nx, ny, dx, jitter, seed);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  })
}

int gsh_gen_sizes(const gsh_gen* g, int32_t* n, int32_t* ne) {
  GSH_API_BODY({
    GSH_REQUIRE(g, GSH_E_ARG, "generator is NULL");
    if (n) *n = (int32_t)g->c.x.size();
    if (ne) *ne = (int32_t)(g->c.tri.size() / 3);
  })
}

int gsh_gen_get(const gsh_gen* g, double* xy, int32_t* elnode, int8_t* open_bnd) {
  GSH_API_BODY({
    GSH_REQUIRE(g, GSH_E_ARG, "generator is NULL");
    const size_t n = g->c.x.size();
    if (xy)
      for (size_t i = 0; i < n; ++i) xy[2 * i] = g->c.x[i], xy[2 * i + 1] = g->c.y[i];
    if (elnode) std::memcpy(elnode, g->c.tri.data(), g->c.tri.size() * sizeof(int32_t));
    if (open_bnd) std::memcpy(open_bnd, g->c.open_bnd.data(), n);
  })
}

int gsh_gen_destroy(gsh_gen* g) {
  delete g;
  return GSH_OK;
}

int gsh_gen_bathymetry(int32_t nx, int32_t ny, double corr_cells, double hmin, double hmax,
                       uint64_t seed, double* depth_out) {
  GSH_API_BODY({
    GSH_REQUIRE(depth_out, GSH_E_ARG, "depth_out is NULL");
    bathymetry(nx, ny, corr_cells, hmin, hmax, seed, depth_out);
  })
}

int gsh_gen_velocity(int32_t npts, const double* xy, int32_t nvrt, const double* sigma, double lx,
                     double ly, double dx, int32_t nmodes, double umax, double shear, uint64_t seed,
                     int32_t ncomp, double* out) {
  GSH_API_BODY({
    GSH_REQUIRE(npts >= 0 && xy && sigma && out && nvrt >= 1, GSH_E_ARG, "velocity: bad pointers");
    GSH_REQUIRE(ncomp == 2 || ncomp == 3, GSH_E_ARG, "velocity: ncomp must be 2 or 3");
    Modes m = make_modes(lx, ly, dx, nmodes, umax, seed);
    for (int32_t p = 0; p < npts; ++p) {
      const double px = xy[2 * p], py = xy[2 * p + 1];
      double u = 0.0, v = 0.0;
      for (int32_t q = 0; q < nmodes; ++q) {
        const double c = std::cos(m.kx[q] * px + m.ky[q] * py + m.ph[q]);
        u += m.amp[q] * m.ky[q] * c;
        v -= m.amp[q] * m.kx[q] * c;
      }
      u *= m.scale;
      v *= m.scale;
      for (int32_t k = 0; k < nvrt; ++k) {
        const double f = 1.0 + shear * sigma[k];
        double* o = out + ((int64_t)p * nvrt + k) * ncomp;
        o[0] = u * f;
        o[1] = v * f;
        if (ncomp == 3) o[2] = 0.0;
      }
    }
  })
}

int gsh_gen_normal(uint64_t seed, int64_t count, double mean, double std_, double* out) {
  GSH_API_BODY({
    GSH_REQUIRE(count >= 0 && (out || count == 0), GSH_E_ARG, "normal: bad arguments");
    Rng rng(seed);
    for (int64_t i = 0; i < count; ++i) out[i] = mean + std_ * rng.normal();
  })
}

}  // extern "C"
