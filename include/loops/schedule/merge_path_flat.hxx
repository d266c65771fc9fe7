/**
 * @file merge_path_flat.hxx
 * @brief Merge-path (flat) schedule: the merge of row ends and nonzeros is
 * cut into tiles of threads_per_block * items_per_thread steps, one tile per
 * workgroup, and each work-item walks items_per_thread consecutive steps
 * (reference include/loops/schedule/merge_path_flat.hxx:92-436).
 *
 * merge_path::preprocess_t precomputes the tile coordinates on the device
 * through the C ABI (loops_merge_path_partition).  Unlike the reference,
 * whose copy constructor drops the precomputed coordinates before the kernel
 * sees them (merge_path_flat.hxx:214-221), the copy here keeps pointing at
 * them, so kernels really read them instead of searching again.  The
 * coordinates are identical either way.
 */
#pragma once

#include <climits>

#include <loops/container/coordinate.hxx>
#include <loops/error.hxx>
#include <loops/schedule.hxx>
#include <loops/stride_ranges.hxx>
#include <loops/util/math.hxx>
#include <loops/util/pair.hxx>
#include <loops/util/search.hxx>
#include <loops_capi.h>

namespace loops {
namespace schedule {

using coord_t = coordinate_t<unsigned int>;

template <typename atoms_type, typename atom_size_type>
class atom_traits<algorithms_t::merge_path_flat, atoms_type, atom_size_type> {
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
class tile_traits<algorithms_t::merge_path_flat, tiles_type, tile_size_type> {
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

namespace merge_path {

/// Host-side preprocess: owns the num_merge_tiles + 1 tile coordinates
/// (merge_path_flat.hxx:150-230).  Copies (including the by-value kernel
/// argument) are non-owning views of the same coordinates.
template <std::size_t THREADS_PER_BLOCK,
          std::size_t ITEMS_PER_THREAD,
          typename tiles_type,
          typename atoms_type,
          typename tile_size_type,
          typename atom_size_type>
class preprocess_t {
 public:
  using tiles_t = tiles_type;
  using tiles_iterator_t = tiles_t*;
  using tile_size_t = tile_size_type;
  using atom_size_t = atom_size_type;
  enum : unsigned int { items_per_tile = THREADS_PER_BLOCK * ITEMS_PER_THREAD };

  /// Computes the coordinates when the search grid would fill the device
  /// (the reference's condition, merge_path_flat.hxx:185-205); otherwise
  /// data() is null and every tile searches for itself.
  preprocess_t(tiles_iterator_t _tiles, tile_size_t _num_tiles,
               atom_size_t _num_atoms, hipStream_t stream = 0)
      : total_work(static_cast<long long>(_num_tiles) + _num_atoms),
        num_merge_tiles(math::ceil_div(total_work, (long long)items_per_tile)),
        d_tile_coordinates(nullptr),
        owner(true) {
    int device = 0, cu_count = 0;
    error::throw_if_exception(hipGetDevice(&device), "hipGetDevice");
    error::throw_if_exception(
        hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount,
                              device),
        "hipDeviceGetAttribute");
    long long grid = math::ceil_div(num_merge_tiles + 1,
                                    (long long)THREADS_PER_BLOCK);
    if (grid >= cu_count && num_merge_tiles > 0) {
      error::throw_if_exception(
          hipMalloc(&d_tile_coordinates,
                    sizeof(coord_t) * (num_merge_tiles + 1)),
          "Tile Coordinates failed allocation.");
      loops_status_t s = loops_merge_path_partition(
          reinterpret_cast<const int32_t*>(_tiles), (int64_t)_num_tiles,
          (int64_t)_num_atoms, (int64_t)items_per_tile,
          reinterpret_cast<loops_coord_t*>(d_tile_coordinates), stream);
      error::throw_if_exception(s != LOOPS_SUCCESS, loops_status_string(s));
    }
  }

  __host__ __device__ preprocess_t(preprocess_t const& rhs)
      : total_work(rhs.total_work),
        num_merge_tiles(rhs.num_merge_tiles),
        d_tile_coordinates(rhs.d_tile_coordinates),
        owner(false) {}

  __host__ __device__ ~preprocess_t() {
#if !defined(__HIP_DEVICE_COMPILE__)
    if (owner && d_tile_coordinates) (void)hipFree(d_tile_coordinates);
#endif
  }

  __host__ __device__ coord_t* data() const { return d_tile_coordinates; }
  __host__ __device__ long long num_tiles() const { return num_merge_tiles; }

 private:
  long long total_work;
  long long num_merge_tiles;
  coord_t* d_tile_coordinates;
  bool owner;
};

}  // namespace merge_path

template <std::size_t THREADS_PER_BLOCK,
          std::size_t ITEMS_PER_THREAD,
          typename tiles_type,
          typename atoms_type,
          typename tile_size_type,
          typename atom_size_type>
class setup<algorithms_t::merge_path_flat, THREADS_PER_BLOCK, ITEMS_PER_THREAD,
            tiles_type, atoms_type, tile_size_type, atom_size_type>
    : public tile_traits<algorithms_t::merge_path_flat, tiles_type,
                         tile_size_type>,
      public atom_traits<algorithms_t::merge_path_flat, atoms_type,
                         atom_size_type> {
 public:
  using tiles_t = tiles_type;
  using atoms_t = atoms_type;
  using tiles_iterator_t = tiles_t*;
  using tile_size_t = tile_size_type;
  using atom_size_t = atom_size_type;
  using meta_t = merge_path::preprocess_t<THREADS_PER_BLOCK, ITEMS_PER_THREAD,
                                          tiles_type, atoms_type,
                                          tile_size_type, atom_size_type>;
  using tile_traits_t =
      tile_traits<algorithms_t::merge_path_flat, tiles_type, tile_size_type>;
  using atom_traits_t =
      atom_traits<algorithms_t::merge_path_flat, atoms_type, atom_size_type>;

  enum : unsigned int {
    threads_per_block = THREADS_PER_BLOCK,
    items_per_thread = ITEMS_PER_THREAD,
    items_per_tile = threads_per_block * items_per_thread,
  };

  /// LDS of the schedule: the tile's two end coordinates and the row-end
  /// offsets of every row the tile touches (merge_path_flat.hxx:288-291).
  struct storage_t {
    coord_t tile_coords[2];
    tiles_t tile_end_offset[items_per_thread + items_per_tile + 1];
  };

  storage_t& buffer;
  meta_t& meta;
  counting_iterator<atoms_t> atoms_counting_it;
  counting_iterator<tiles_t> tiles_counting_it;

  __device__ __forceinline__ setup(meta_t& _meta, storage_t& _buffer,
                                   tiles_iterator_t _tiles,
                                   tile_size_t _num_tiles,
                                   atom_size_t _num_atoms)
      : tile_traits_t(_num_tiles, _tiles),
        atom_traits_t(_num_atoms),
        buffer(_buffer),
        meta(_meta),
        total_work(static_cast<long long>(_num_tiles) + _num_atoms),
        num_merge_tiles(math::ceil_div(total_work, (long long)items_per_tile)),
        tile_num_tiles(0),
        tile_num_atoms(0) {}

  /// Tile coordinates (precomputed or searched by two lanes), LDS gather of
  /// the tile's row ends, then each lane's start by a search inside LDS
  /// (merge_path_flat.hxx:321-385).  Returns the lane's coordinate relative
  /// to the tile start, or {UINT_MAX, UINT_MAX} for a block past the last
  /// tile (block-uniform).
  __device__ __forceinline__ coord_t init() {
    long long tid = (long long)blockIdx.x * gridDim.y + blockIdx.y;
    if (tid >= num_merge_tiles) return coord_t{UINT_MAX, UINT_MAX};

    if (threadIdx.x < 2) {
      if (meta.data() == nullptr) {
        long long diagonal = (tid + threadIdx.x) * (long long)items_per_tile;
        buffer.tile_coords[threadIdx.x] = search::_binary_search<long long>(
            diagonal, tile_traits_t::begin() + 1, 0LL,
            (long long)tile_traits_t::size(), (long long)atom_traits_t::size());
      } else {
        buffer.tile_coords[threadIdx.x] = meta.data()[tid + threadIdx.x];
      }
    }
    __syncthreads();

    coord_t tile_start = buffer.tile_coords[0];
    coord_t tile_end = buffer.tile_coords[1];
    tile_num_tiles = tile_end.x - tile_start.x;
    tile_num_atoms = tile_end.y - tile_start.y;

    const tiles_t* end_offsets = tile_traits_t::begin() + 1;
    long long last_row = (long long)tile_traits_t::size() - 1;
    for (long long item = threadIdx.x;
         item < (long long)tile_num_tiles + items_per_thread;
         item += threads_per_block) {
      long long r = (long long)tile_start.x + item;
      buffer.tile_end_offset[item] = end_offsets[r < last_row ? r : last_row];
    }
    tiles_counting_it = counting_iterator<tiles_t>(tiles_t(tile_start.x));
    atoms_counting_it = counting_iterator<atoms_t>(atoms_t(tile_start.y));
    __syncthreads();

    coord_t thread_start = search::_binary_search<long long>(
        (long long)threadIdx.x * items_per_thread, buffer.tile_end_offset,
        (long long)tile_start.y, (long long)tile_num_tiles,
        (long long)tile_num_atoms);
    __syncthreads();
    return thread_start;
  }

  __device__ __forceinline__ bool is_valid_accessor(coord_t& coord) const {
    return coord.x != UINT_MAX && coord.y != UINT_MAX;
  }

  /// 0 .. items_per_thread - 1.
  __device__ __forceinline__ step_range_t<int> virtual_idx() const {
    return custom_stride_range(int(0), int(items_per_thread), int(1));
  }

  /// Nonzero under the coordinate, clamped to the last one (:408-410).
  __device__ __forceinline__ atoms_t atom_idx(int, coord_t& coord) const {
    long long a = (long long)atoms_counting_it[coord.y];
    long long last = (long long)atom_traits_t::size() - 1;
    return atoms_t(a < last ? a : last);
  }

  /// Row under the coordinate (:418-420).
  __device__ __forceinline__ tiles_t tile_idx(coord_t& coord) const {
    return tiles_counting_it[coord.x];
  }

  __device__ __forceinline__ tile_size_t num_tiles() const {
    return tile_num_tiles;
  }
  __device__ __forceinline__ atom_size_t num_atoms() const {
    return tile_num_atoms;
  }

 private:
  long long total_work;
  long long num_merge_tiles;
  tile_size_t tile_num_tiles;
  atom_size_t tile_num_atoms;
};

}  // namespace schedule
}  // namespace loops
