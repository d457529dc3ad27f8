// ssor_layout.cpp -- smallest-last colouring and the colour-major split CSR of the SSOR solver.
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
  L.lptr.assign(np + 1, 0);
  L.uptr.assign(np + 1, 0);
  L.dsrc.assign(np, -1);
  std::vector<std::pair<int32_t, int32_t>> lo, up;  // (new column, slot)
  for (int32_t v = 0; v < np; ++v) {
    const int32_t i = L.perm[v];
    lo.clear();
    up.clear();
    if (i >= 0) {
      const int32_t c = L.colour[i];
      for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k) {
        const int32_t j = indices[k];
        if (j == i) {
          L.dsrc[v] = k;
          continue;
        }
        const int32_t w = L.newpos[j];
        (w < L.cstart[c] ? lo : up).push_back({w, k});
      }
      GSH_REQUIRE(L.dsrc[v] >= 0, GSH_E_ARG, "SSOR: every row needs a diagonal entry in the pattern");
      std::sort(lo.begin(), lo.end());
      std::sort(up.begin(), up.end());
    }
    for (const auto& e : lo) L.lcol.push_back(e.first), L.lsrc.push_back(e.second);
    for (const auto& e : up) L.ucol.push_back(e.first), L.usrc.push_back(e.second);
    L.lptr[v + 1] = (int32_t)L.lcol.size();
    L.uptr[v + 1] = (int32_t)L.ucol.size();
  }
}

}  // namespace gsh
