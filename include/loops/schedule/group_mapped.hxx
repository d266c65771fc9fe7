/**
 * @file group_mapped.hxx
 * @brief Group-mapped schedule: a group of threads_per_tile work-items owns
 * threads_per_tile consecutive tiles (rows) and shares out their atoms
 * (nonzeros) evenly (reference include/loops/schedule/group_mapped.hxx:
 * 80-248).
 *
 * CDNA notes: groups of up to 64 lanes live inside one wavefront and scan
 * with wave shuffles.  Groups of several wavefronts (a multiple of 64 lanes
 * dividing the block, block_mapped included) scan each wavefront with
 * shuffles and carry the wavefront totals through LDS; their sync() is a
 * workgroup barrier, so, as with the reference's block-sized groups, every
 * lane of the block must reach it (atom_accessor is called by all lanes).
 * warp_mapped keeps the reference's 32 lanes (group_mapped.hxx:239-241), so
 * its row-to-group partition is the reference's; wave_mapped is the 64-lane
 * wavefront of CDNA.
 */
#pragma once

#include <loops/schedule.hxx>
#include <loops/stride_ranges.hxx>

namespace loops {
namespace schedule {

/// Thread group of `tile_size` lanes inside a block of `block_size` lanes
/// (stands in for the reference's cooperative_groups tiled_partition).
template <std::size_t block_size, std::size_t tile_size>
struct thread_group_t {
  static_assert(tile_size == block_size ||
                    (tile_size <= 64 && (tile_size & (tile_size - 1)) == 0) ||
                    (tile_size % 64 == 0 && block_size % tile_size == 0),
                "group size must be a power of two <= 64, a multiple of 64 "
                "dividing the block size, or the block size");
  static_assert(block_size % 64 == 0 || block_size == tile_size,
                "block size must be whole wavefronts");
  __device__ unsigned int thread_rank() const {
    return threadIdx.x % tile_size;
  }
  __device__ unsigned int meta_group_rank() const {
    return threadIdx.x / tile_size;
  }
  __device__ static constexpr unsigned int size() { return tile_size; }
  __device__ static constexpr unsigned int num_threads() { return tile_size; }
  __device__ void sync() const {
    if constexpr (tile_size > 64) {
      __syncthreads();
    } else {
      // Lanes of one wavefront: order their LDS traffic.
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
};

template <typename atoms_type, typename atom_size_type>
class atom_traits<algorithms_t::group_mapped, atoms_type, atom_size_type> {
 public:
  using atoms_t = atoms_type;
  using atoms_iterator_t = atoms_t*;
  using atom_size_t = atom_size_type;
  __host__ __device__ atom_traits() : size_(0), atoms_(nullptr) {}
  __host__ __device__ atom_traits(atom_size_t size)
      : size_(size), atoms_(nullptr) {}
  __host__ __device__ atom_traits(atom_size_t size, atoms_iterator_t atoms)
      : size_(size), atoms_(atoms) {}
  __host__ __device__ atom_size_t size() const { return size_; }
  __host__ __device__ atoms_iterator_t begin() { return atoms_; }
  __host__ __device__ atoms_iterator_t end() { return atoms_ + size_; }

 private:
  atom_size_t size_;
  atoms_iterator_t atoms_;
};

template <typename tiles_type, typename tile_size_type>
class tile_traits<algorithms_t::group_mapped, tiles_type, tile_size_type> {
 public:
  using tiles_t = tiles_type;
  using tiles_iterator_t = tiles_t*;
  using tile_size_t = tile_size_type;
  __host__ __device__ tile_traits() : size_(0), tiles_(nullptr) {}
  __host__ __device__ tile_traits(tile_size_t size, tiles_iterator_t tiles)
      : size_(size), tiles_(tiles) {}
  __host__ __device__ tile_size_t size() const { return size_; }
  __host__ __device__ tiles_iterator_t begin() const { return tiles_; }
  __host__ __device__ tiles_iterator_t end() const { return tiles_ + size_; }

 private:
  tile_size_t size_;
  tiles_iterator_t tiles_;
};

template <std::size_t THREADS_PER_BLOCK,
          std::size_t THREADS_PER_TILE,
          typename tiles_type,
          typename atoms_type,
          typename tile_size_type,
          typename atom_size_type>
class setup<algorithms_t::group_mapped, THREADS_PER_BLOCK, THREADS_PER_TILE,
            tiles_type, atoms_type, tile_size_type, atom_size_type>
    : public tile_traits<algorithms_t::group_mapped, tiles_type,
                         tile_size_type>,
      public atom_traits<algorithms_t::group_mapped, atoms_type,
                         atom_size_type> {
 public:
  using tiles_t = tiles_type;
  using atoms_t = atoms_type;
  using tile_size_t = tile_size_type;
  using atom_size_t = atom_size_type;
  using tile_traits_t =
      tile_traits<algorithms_t::group_mapped, tiles_type, tile_size_type>;
  using atom_traits_t =
      atom_traits<algorithms_t::group_mapped, atoms_type, atom_size_type>;
  using partition_t = thread_group_t<THREADS_PER_BLOCK, THREADS_PER_TILE>;

  enum : unsigned int {
    threads_per_block = THREADS_PER_BLOCK,
    threads_per_tile = THREADS_PER_TILE,
    tiles_per_block = THREADS_PER_BLOCK / THREADS_PER_TILE,
    waves_per_block = (THREADS_PER_BLOCK + 63) / 64,
  };

  /// LDS owned by the schedule (declare it __shared__ in the kernel).
  struct __align__(16) storage_t {
    atoms_t tile_aggregates[tiles_per_block];
    atoms_t atoms_offsets[threads_per_block];
    tiles_t tiles_indices[threads_per_block];
    atoms_t wave_totals[waves_per_block];
  };

  storage_t& buffer;

  __device__ __forceinline__ setup(storage_t& _buffer, tiles_t* _tiles,
                                   tile_size_t _num_tiles,
                                   atom_size_t _num_atoms)
      : tile_traits_t(_num_tiles, _tiles),
        atom_traits_t(_num_atoms),
        buffer(_buffer) {}

  /// Records which tile (row) each lane of each group stands for
  /// (group_mapped.hxx:146-159).
  __device__ __forceinline__ partition_t partition() {
    partition_t p;
    long long index = global_rank();
    unsigned int tile_index = p.thread_rank() + p.meta_group_rank() * p.size();
    buffer.tiles_indices[tile_index] =
        index < (long long)tile_traits_t::size() ? tiles_t(index) : tiles_t(-1);
    return p;
  }

  /// Exclusive scan of the group's row lengths into LDS; returns the atoms
  /// this lane visits: rank, rank + size, ... < group aggregate
  /// (group_mapped.hxx:161-186).  Every lane of the block must call it.
  __device__ step_range_t<atoms_t> atom_accessor(partition_t& p) {
    atoms_t* p_st = buffer.atoms_offsets + p.meta_group_rank() * threads_per_tile;
    long long index = global_rank();
    atoms_t num_atoms = 0;
    if (index < (long long)tile_traits_t::size())
      num_atoms = tile_traits_t::begin()[index + 1] - tile_traits_t::begin()[index];
    atoms_t excl = exclusive_scan(p, num_atoms);
    p_st[p.thread_rank()] = excl;
    if (p.thread_rank() == p.size() - 1)
      buffer.tile_aggregates[p.meta_group_rank()] = excl + num_atoms;
    p.sync();
    atoms_t aggregate = buffer.tile_aggregates[p.meta_group_rank()];
    return custom_stride_range(atoms_t(p.thread_rank()), aggregate,
                               atoms_t(p.size()));
  }

  /// Number of valid tiles in this lane's group (group_mapped.hxx:188-201).
  __device__ __forceinline__ int get_length(partition_t& p) const {
    long long first = global_rank() - p.thread_rank();
    long long n = (long long)tile_traits_t::size() - first;
    if (n > (long long)p.size()) n = p.size();
    if (n < 0) n = 0;
    return int(n);
  }

  /// Group-local tile holding virtual atom v: upper_bound over the scanned
  /// offsets, minus one (group_mapped.hxx:203-213).
  __device__ __forceinline__ tiles_t tile_accessor(atoms_t& virtual_atom,
                                                   partition_t& p) {
    int length = get_length(p);
    const atoms_t* p_st =
        buffer.atoms_offsets + p.meta_group_rank() * threads_per_tile;
    int lo = 0, hi = length;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (p_st[mid] <= virtual_atom) lo = mid + 1;
      else hi = mid;
    }
    return tiles_t(lo - 1);
  }

  __device__ __forceinline__ bool is_valid_accessor(tiles_t& tile_id,
                                                    partition_t& p) {
    return tile_id < get_length(p);
  }

  __device__ __forceinline__ tiles_t tile_id(tiles_t& v_tile_id,
                                             partition_t& p) {
    return buffer.tiles_indices[v_tile_id + p.meta_group_rank() * p.size()];
  }

  __device__ __forceinline__ atoms_t atom_id(atoms_t& v_atom, tiles_t& tile_id,
                                             tiles_t& v_tile_id,
                                             partition_t& p) {
    const atoms_t* p_st =
        buffer.atoms_offsets + p.meta_group_rank() * threads_per_tile;
    return tile_traits_t::begin()[tile_id] + v_atom - p_st[v_tile_id];
  }

 private:
  __device__ static long long global_rank() {
    return (long long)blockIdx.x * blockDim.x + threadIdx.x;
  }

  /// Exclusive prefix sum over the group (integer, exact).  A group of
  /// several wavefronts adds the totals of its earlier wavefronts, read
  /// from LDS between two workgroup barriers.
  __device__ atoms_t exclusive_scan(partition_t& p, atoms_t v) {
    constexpr int width = threads_per_tile < 64 ? threads_per_tile : 64;
    const int lane_in = threadIdx.x % width;
    atoms_t inc = v;
#pragma unroll
    for (int d = 1; d < width; d <<= 1) {
      atoms_t up = __shfl_up(inc, d, width);
      if (lane_in >= d) inc += up;
    }
    if constexpr (threads_per_tile <= 64) {
      return inc - v;
    } else {
      constexpr int waves_per_group = threads_per_tile / 64;
      const int wave = threadIdx.x / 64;
      const int first_wave = wave - wave % waves_per_group;
      if ((threadIdx.x & 63) == 63) buffer.wave_totals[wave] = inc;
      __syncthreads();
      atoms_t carry = 0;
      for (int w = first_wave; w < wave; ++w) carry += buffer.wave_totals[w];
      __syncthreads();
      return carry + inc - v;
    }
  }
};

template <std::size_t threads_per_block, typename tiles_t, typename atoms_t>
using warp_mapped =
    setup<algorithms_t::group_mapped, threads_per_block, 32, tiles_t, atoms_t>;

template <std::size_t threads_per_block, typename tiles_t, typename atoms_t>
using wave_mapped =
    setup<algorithms_t::group_mapped, threads_per_block, 64, tiles_t, atoms_t>;

template <std::size_t threads_per_block, typename tiles_t, typename atoms_t>
using block_mapped = setup<algorithms_t::group_mapped, threads_per_block,
                           threads_per_block, tiles_t, atoms_t>;

}  // namespace schedule
}  // namespace loops
