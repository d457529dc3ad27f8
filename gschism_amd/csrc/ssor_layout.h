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
// identity rows, and the off-diagonal part split into the strictly lower part (columns of lower
// colours) and the strictly upper part, each row's entries in ascending new column. Both parts
// are stored ELL per colour: colour c keeps a width w (its longest row part) and entry m of the
// row at offset r (= new row - cstart[c]) sits at off[c] + m * rows[c] + r, so a wave reads 64
// rows' entry m with one coalesced load. Unused entries have column -1 and slot -1; `*src` is
// the slot of each entry in the original CSR value array.
struct SsorLayout {
  int32_t n = 0, np = 0, ncol = 0, tile_rows = 0;
  std::vector<int32_t> colour;               // [n]
  std::vector<int32_t> cstart;               // [ncol + 1] first padded row of each colour
  std::vector<int32_t> perm;                 // [np] new row -> original row, -1 for padding
  std::vector<int32_t> newpos;               // [n] original row -> new row
  std::vector<int32_t> dsrc;                 // [np] slot of the diagonal, -1 for padding
  std::vector<int32_t> wl, wu;               // [ncol] ELL widths
  std::vector<int64_t> offl, offu;           // [ncol] ELL offsets
  std::vector<int32_t> lcol, lsrc, ucol, usrc;  // ELL entries
  std::vector<int32_t> tile_colour;          // [np / tile_rows]
};

// Throws gsh::Error(GSH_E_ARG) when a row has no diagonal entry.
void build_ssor_layout(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t tile_rows, SsorLayout& L);

}  // namespace gsh
