// ssor_layout.h -- multicolour ordering of a CSR pattern for the SSOR-preconditioned solver.
#pragma once

#include <cstdint>
#include <vector>

namespace gsh {

// Greedy first-fit colouring of the graph of a CSR pattern (diagonal ignored) in
// smallest-last order: vertices are removed by least (remaining degree, index), then coloured
// in reverse removal order. Returns the number of colours.
int32_t colour_smallest_last(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* colour);

// Colour-major row order with every colour block padded to a multiple of `tile_rows` with
// identity rows, and the matrix split into strictly-lower (lower colours) and strictly-upper
// (higher colours) parts, each row sorted by new column index. The *src arrays give the slot
// of each entry in the original CSR value array.
struct SsorLayout {
  int32_t n = 0, np = 0, ncol = 0, tile_rows = 0;
  std::vector<int32_t> colour;               // [n]
  std::vector<int32_t> cstart;               // [ncol + 1] first padded row of each colour
  std::vector<int32_t> perm;                 // [np] new row -> original row, -1 for padding
  std::vector<int32_t> newpos;               // [n] original row -> new row
  std::vector<int32_t> lptr, lcol, lsrc;     // lower part
  std::vector<int32_t> uptr, ucol, usrc;     // upper part
  std::vector<int32_t> dsrc;                 // [np] slot of the diagonal, -1 for padding
};

// Throws gsh::Error(GSH_E_ARG) when a row has no diagonal entry.
void build_ssor_layout(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t tile_rows, SsorLayout& L);

}  // namespace gsh
