// ssor_layout.cpp -- smallest-last colouring and the colour-major split ELL of the SSOR solver.
#include "ssor_layout.h"

#include <algorithm>
#include <queue>
#include <utility>

#include "abi_util.h"

namespace gsh {

int32_t colour_smallest_last(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* colour) {
  if (n <= 0) return 0;
  std::vector<int32_t> deg(n, 0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) deg[i] += indices[k] != i;
  // min-queue on (degree, index); stale entries are skipped when popped
  using Key = std::pair<int32_t, int32_t>;
  std::priority_queue<Key, std::vector<Key>, std::greater<Key>> q;
  for (int32_t i = 0; i < n; ++i) q.push({deg[i], i});
  std::vector<char> removed(n, 0);
  std::vector<int32_t> order;
  order.reserve(n);
  while (!q.empty()) {
    const auto [d, i] = q.top();
    q.pop();
    if (removed[i] || d != deg[i]) continue;
    removed[i] = 1;
    order.push_back(i);
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      const int32_t j = indices[k];
      if (j != i && !removed[j]) q.push({--deg[j], j});
    }
  }
  std::fill(colour, colour + n, -1);
  std::vector<int32_t> seen(n + 2, -1);
  int32_t ncol = 0;
  for (int32_t m = n - 1; m >= 0; --m) {
    const int32_t i = order[m];
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      const int32_t j = indices[k];
      if (j != i && colour[j] >= 0) seen[colour[j]] = m;
    }
    int32_t c = 0;
    while (seen[c] == m) ++c;
    colour[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

void build_ssor_layout(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t tile_rows, SsorLayout& L) {
  L = SsorLayout{};
  L.n = n;
  L.tile_rows = tile_rows;
  L.colour.assign(n, 0);
  L.ncol = colour_smallest_last(n, indptr, indices, L.colour.data());
  const int32_t C = L.ncol;
  std::vector<int64_t> count(C, 0);
  for (int32_t i = 0; i < n; ++i) count[L.colour[i]]++;
  L.cstart.assign(C + 1, 0);
  int64_t acc = 0;
  for (int32_t c = 0; c < C; ++c) {
    acc += (count[c] + tile_rows - 1) / tile_rows * tile_rows;
    GSH_REQUIRE(acc < INT32_MAX, GSH_E_ARG, "SSOR layout: padded size exceeds int32");
    L.cstart[c + 1] = (int32_t)acc;
  }
  const int32_t np = L.cstart[C];
  L.np = np;
  L.perm.assign(np, -1);
  L.newpos.assign(n, 0);
  std::vector<int32_t> fill(C, 0);
  for (int32_t i = 0; i < n; ++i) {
    const int32_t c = L.colour[i];
    const int32_t v = L.cstart[c] + fill[c]++;
    L.perm[v] = i;
    L.newpos[i] = v;
  }
  L.dsrc.assign(np, -1);
  // sorted lower / upper lists of every row: (new column, slot)
  std::vector<std::vector<std::pair<int32_t, int32_t>>> lo(np), up(np);
  for (int32_t v = 0; v < np; ++v) {
    const int32_t i = L.perm[v];
    if (i < 0) continue;
    const int32_t c = L.colour[i];
    for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      const int32_t j = indices[k];
      if (j == i) {
        L.dsrc[v] = k;
        continue;
      }
      const int32_t w = L.newpos[j];
      (w < L.cstart[c] ? lo[v] : up[v]).push_back({w, k});
    }
    GSH_REQUIRE(L.dsrc[v] >= 0, GSH_E_ARG, "SSOR: every row needs a diagonal entry in the pattern");
    std::sort(lo[v].begin(), lo[v].end());
    std::sort(up[v].begin(), up[v].end());
  }
  L.wl.assign(C, 0);
  L.wu.assign(C, 0);
  L.offl.assign(C, 0);
  L.offu.assign(C, 0);
  int64_t tl = 0, tu = 0;
  for (int32_t c = 0; c < C; ++c) {
    for (int32_t v = L.cstart[c]; v < L.cstart[c + 1]; ++v) {
      L.wl[c] = std::max<int32_t>(L.wl[c], (int32_t)lo[v].size());
      L.wu[c] = std::max<int32_t>(L.wu[c], (int32_t)up[v].size());
    }
    const int64_t rows = L.cstart[c + 1] - L.cstart[c];
    L.offl[c] = tl;
    L.offu[c] = tu;
    tl += rows * L.wl[c];
    tu += rows * L.wu[c];
  }
  GSH_REQUIRE(tl < INT32_MAX && tu < INT32_MAX, GSH_E_ARG, "SSOR layout: ELL size exceeds int32");
  L.lcol.assign(tl, -1);
  L.lsrc.assign(tl, -1);
  L.ucol.assign(tu, -1);
  L.usrc.assign(tu, -1);
  for (int32_t c = 0; c < C; ++c) {
    const int64_t rows = L.cstart[c + 1] - L.cstart[c];
    for (int32_t v = L.cstart[c]; v < L.cstart[c + 1]; ++v) {
      const int64_t r = v - L.cstart[c];
      for (size_t m = 0; m < lo[v].size(); ++m) {
        L.lcol[L.offl[c] + m * rows + r] = lo[v][m].first;
        L.lsrc[L.offl[c] + m * rows + r] = lo[v][m].second;
      }
      for (size_t m = 0; m < up[v].size(); ++m) {
        L.ucol[L.offu[c] + m * rows + r] = up[v][m].first;
        L.usrc[L.offu[c] + m * rows + r] = up[v][m].second;
      }
    }
  }
  L.tile_colour.assign(np / tile_rows, 0);
  for (int32_t c = 0; c < C; ++c)
    for (int32_t t = L.cstart[c] / tile_rows; t < L.cstart[c + 1] / tile_rows; ++t) L.tile_colour[t] = c;
}

}  // namespace gsh
