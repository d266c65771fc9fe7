/**
 * @file coordinate.hxx
 * @brief (x, y) coordinate of the merge path (reference
 * include/loops/container/coordinate.hxx:14-18).  x indexes rows (tiles),
 * y indexes nonzeros (atoms).
 */
#pragma once

namespace loops {
template <typename index_t>
struct coordinate_t {
  index_t x;
  index_t y;
};
}  // namespace loops
