/**
 * @file work_oriented.hxx
 * @brief Work-oriented schedule: every work-item of the grid receives an
 * equal, contiguous share of the merge path (rows + nonzeros) and finds its
 * start and end with two merge-path searches (reference
 * include/loops/schedule/work_oriented.hxx:97-237).
 *
 * init() returns {start, end} where each is {tile (row), atom (nonzero)};
 * tiles(map) yields the rows whose end falls in the share, atoms(row, map)
 * the nonzeros of that row from the share's current position (advancing
 * it), and remainder_tiles/remainder_atoms the partial last row.
 */
#pragma once

#include <loops/schedule.hxx>
#include <loops/stride_ranges.hxx>
#include <loops/util/math.hxx>
#include <loops/util/pair.hxx>
#include <loops/util/search.hxx>

namespace loops {
namespace schedule {

template <typename atoms_type, typename atom_size_type>
class atom_traits<algorithms_t::work_oriented, atoms_type, atom_size_type> {
 public:
  using atoms_t = atoms_type;
  using atom_size_t = atom_size_type;
  __host__ __device__ atom_traits() : size_(0) {}
  __host__ __device__ atom_traits(atom_size_t size) : size_(size) {}
  __host__ __device__ atom_size_t size() const { return size_; }

 private:
  atom_size_t size_;
};

template <typename tiles_type, typename tile_size_type>
class tile_traits<algorithms_t::work_oriented, tiles_type, tile_size_type> {
 public:
  using tiles_t = tiles_type;
  using tiles_iterator_t = tiles_t*;
  using tile_size_t = tile_size_type;
  __host__ __device__ tile_traits() : size_(0), tiles_(nullptr) {}
  __host__ __device__ tile_traits(tile_size_t size, tiles_iterator_t tiles)
      : size_(size), tiles_(tiles) {}
  __host__ __device__ tile_size_t size() const { return size_; }
  __host__ __device__ tiles_iterator_t begin() const { return tiles_; }

 private:
  tile_size_t size_;
  tiles_iterator_t tiles_;
};

template <std::size_t THREADS_PER_BLOCK,
          std::size_t ITEMS_PER_THREAD,
          typename tiles_type,
          typename atoms_type,
          typename tile_size_type,
          typename atom_size_type>
class setup<algorithms_t::work_oriented, THREADS_PER_BLOCK, ITEMS_PER_THREAD,
            tiles_type, atoms_type, tile_size_type, atom_size_type>
    : public tile_traits<algorithms_t::work_oriented, tiles_type,
                         tile_size_type>,
      public atom_traits<algorithms_t::work_oriented, atoms_type,
                         atom_size_type> {
 public:
  using tiles_t = tiles_type;
  using atoms_t = atoms_type;
  using tile_size_t = tile_size_type;
  using atom_size_t = atom_size_type;
  using tile_traits_t =
      tile_traits<algorithms_t::work_oriented, tiles_type, tile_size_type>;
  using atom_traits_t =
      atom_traits<algorithms_t::work_oriented, atoms_type, atom_size_type>;
  using coord_pair_t = pair_t<atoms_t, atoms_t>;   // {tile, atom}
  using map_t = pair_t<coord_pair_t, coord_pair_t>; // {start, end}

  enum : unsigned int {
    threads_per_block = THREADS_PER_BLOCK,
    items_per_thread = ITEMS_PER_THREAD,
    items_per_tile = threads_per_block * items_per_thread,
  };

  __device__ __forceinline__ setup(tiles_t* _tiles, tile_size_t _num_tiles,
                                   atom_size_t _num_atoms)
      : tile_traits_t(_num_tiles, _tiles),
        atom_traits_t(_num_atoms),
        total_work(static_cast<long long>(_num_tiles) + _num_atoms),
        num_threads(static_cast<long long>(gridDim.x) * threads_per_block),
        work_per_thread(math::ceil_div(total_work, num_threads)) {}

  /// Start and end coordinates of this work-item's share
  /// (work_oriented.hxx:150-167).
  __device__ __forceinline__ map_t init() const {
    long long tid = threadIdx.x + static_cast<long long>(blockIdx.x) * blockDim.x;
    long long upper = work_per_thread * tid;
    if (upper > total_work) upper = total_work;
    long long lower = upper + work_per_thread;
    if (lower > total_work) lower = total_work;
    auto st = search(upper);
    auto en = search(lower);
    return map_t{coord_pair_t{atoms_t(st.x), atoms_t(st.y)},
                 coord_pair_t{atoms_t(en.x), atoms_t(en.y)}};
  }

  __device__ __forceinline__ step_range_t<tiles_t> tiles(map_t& m) const {
    return custom_stride_range(tiles_t(m.first.first), tiles_t(m.second.first),
                               tiles_t(1));
  }

  __device__ __forceinline__ step_range_t<atoms_t> atoms(tiles_t t,
                                                         map_t& m) const {
    atoms_t nz_next_start = tile_traits_t::begin()[t + 1];
    atoms_t nz_start = m.first.second;
    m.first.second += (nz_next_start - nz_start);
    return custom_stride_range(nz_start, nz_next_start, atoms_t(1));
  }

  __device__ __forceinline__ step_range_t<tiles_t> remainder_tiles(
      map_t& m) const {
    return custom_stride_range(tiles_t(m.second.first),
                               tiles_t(m.second.first + 1), tiles_t(1));
  }

  __device__ __forceinline__ step_range_t<atoms_t> remainder_atoms(
      map_t& m) const {
    return custom_stride_range(atoms_t(m.first.second),
                               atoms_t(m.second.second), atoms_t(1));
  }

  __device__ __forceinline__ long long work_per_thread_count() const {
    return work_per_thread;
  }

 private:
  __device__ __forceinline__ coordinate_t<unsigned int> search(
      long long diagonal) const {
    return search::_binary_search<long long>(
        diagonal, tile_traits_t::begin() + 1, 0LL,
        static_cast<long long>(tile_traits_t::size()),
        static_cast<long long>(atom_traits_t::size()));
  }

  long long total_work;
  long long num_threads;
  long long work_per_thread;
};

}  // namespace schedule
}  // namespace loops
