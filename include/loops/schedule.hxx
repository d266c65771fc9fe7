/**
 * @file schedule.hxx
 * @brief The load-balancing schedule API: loops::schedule::setup<algorithm,
 * threads_per_block, threads_per_tile (or items_per_thread), tiles_t,
 * atoms_t, tile_size_t, atom_size_t> (reference include/loops/schedule.hxx:
 * 23-63).  "Tiles" are rows, "atoms" are nonzeros.
 *
 * Written for CDNA: wavefronts are 64 lanes.  warp_mapped keeps the
 * reference's 32-lane groups; wave_mapped is the 64-lane wavefront; group
 * scans use wave shuffles within 64 lanes.
 */
#pragma once

#include <cstddef>

namespace loops {
namespace schedule {

enum algorithms_t {
  merge_path_flat,  ///< even share of (rows + nonzeros) per thread / tile
  work_oriented,    ///< merge path with one contiguous range per thread
  thread_mapped,    ///< one row per thread
  group_mapped,     ///< a group (wave or block) per set of rows
  bucketing,        ///< declared by the reference, not implemented there
};

template <algorithms_t scheme, typename atoms_t, typename atom_size_t>
class atom_traits;

template <algorithms_t scheme, typename tiles_t, typename tile_size_t>
class tile_traits;

template <algorithms_t scheme,
          std::size_t threads_per_block,
          std::size_t threads_per_tile,
          typename tiles_t,
          typename atoms_t,
          typename tile_size_t = std::size_t,
          typename atom_size_t = std::size_t>
class setup;

}  // namespace schedule
}  // namespace loops

#include <loops/schedule/thread_mapped.hxx>
#include <loops/schedule/group_mapped.hxx>
#include <loops/schedule/work_oriented.hxx>
#include <loops/schedule/merge_path_flat.hxx>
