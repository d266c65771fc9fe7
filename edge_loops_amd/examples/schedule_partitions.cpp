// Dumps the partitions that the device schedule API computes, so they can be
// compared bit for bit with the oracle's restatement of the reference's
// partitions.  Usage:
//   loops.schedule_partitions -m matrix.mtx -o outdir
// writes raw little-endian arrays:
//   mp_<TPB>_<IPT>.bin       uint32 [tiles][TPB][2]: the global start
//   mp_<TPB>_<IPT>_meta.bin  coordinate (row, nonzero) of every lane of every
//                            tile after setup<merge_path_flat>::init(), with
//                            the tile coordinates searched in the kernel and
//                            read from preprocess_t (merge_path_flat.hxx:
//                            321-385), for (128, 5) and (64, 3), the
//                            reference's tiles for 4- and 8-byte values
//                            (merge_path_flat.cuh:102-103);
//   wo_<TPB>.bin             uint32 [threads][4]: the (start, end)
//                            coordinates setup<work_oriented>::init() gives
//                            every work-item of a 2 x CU grid
//                            (work_oriented.hxx:150-167);
//   gm_<TPB>_<TPT>.bin       int64 [nnz][2]: for every nonzero, the group
//                            (blockIdx * TPB / TPT + group in block) and the
//                            virtual atom under which setup<group_mapped>'s
//                            accessors reach it (group_mapped.hxx:146-235),
//                            for block_mapped<128>, warp_mapped<256> (32
//                            lanes), wave_mapped<256> (64 lanes), 128-lane
//                            groups in 256-lane blocks and 8-lane groups.
// and prints "wo_threads <n>" and "tiles_<TPB>_<IPT> <n>".
#include <hip/hip_runtime.h>

#include <cstdint>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include <loops/container/csr.hxx>
#include <loops/container/market.hxx>
#include <loops/error.hxx>
#include <loops/schedule.hxx>
#include <loops/util/device.hxx>
#include <loops/util/launch.hxx>
#include <loops/util/math.hxx>

using namespace loops;

template <std::size_t TPB, std::size_t IPT, typename meta_t>
__global__ void __launch_bounds__(TPB)
    k_mp_coords(meta_t meta, std::size_t rows, std::size_t nnz, int* offsets,
                unsigned* out) {
  using setup_t = schedule::setup<schedule::algorithms_t::merge_path_flat, TPB, IPT,
                                  int, int, std::size_t, std::size_t>;
  __shared__ typename setup_t::storage_t temporary_storage;
  setup_t config(meta, temporary_storage, offsets, rows, nnz);
  auto map = config.init();
  if (!config.is_valid_accessor(map)) return;
  const std::size_t slot = ((std::size_t)blockIdx.x * TPB + threadIdx.x) * 2;
  out[slot] = (unsigned)config.tile_idx(map);
  out[slot + 1] = (unsigned)config.atoms_counting_it[map.y];
}

template <std::size_t TPB>
__global__ void __launch_bounds__(TPB)
    k_wo_coords(std::size_t rows, std::size_t nnz, int* offsets, unsigned* out) {
  using setup_t = schedule::setup<schedule::algorithms_t::work_oriented, TPB, 1, int, int,
                                  std::size_t, std::size_t>;
  setup_t config(offsets, rows, nnz);
  auto map = config.init();
  const std::size_t tid = (std::size_t)blockIdx.x * TPB + threadIdx.x;
  out[4 * tid + 0] = (unsigned)map.first.first;
  out[4 * tid + 1] = (unsigned)map.first.second;
  out[4 * tid + 2] = (unsigned)map.second.first;
  out[4 * tid + 3] = (unsigned)map.second.second;
}

template <std::size_t TPB, std::size_t TPT>
__global__ void __launch_bounds__(TPB)
    k_gm_map(std::size_t rows, std::size_t nnz, int* offsets, long long* out) {
  using setup_t = schedule::setup<schedule::algorithms_t::group_mapped, TPB, TPT, int, int>;
  __shared__ typename setup_t::storage_t temporary_storage;
  setup_t config(temporary_storage, offsets, rows, nnz);
  auto p = config.partition();
  const long long group = (long long)blockIdx.x * (TPB / TPT) + p.meta_group_rank();
  for (auto virtual_atom : config.atom_accessor(p)) {
    auto virtual_tile = config.tile_accessor(virtual_atom, p);
    if (!(config.is_valid_accessor(virtual_tile, p))) continue;
    auto row = config.tile_id(virtual_tile, p);
    auto nz = config.atom_id(virtual_atom, row, virtual_tile, p);
    out[2 * (long long)nz] = group;
    out[2 * (long long)nz + 1] = virtual_atom;
  }
}

template <typename T>
static void write_bin(const std::string& path, const T* d, std::size_t count) {
  std::vector<T> h(count);
  if (count)
    error::throw_if_exception(hipMemcpy(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost),
                              "hipMemcpy");
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(h.data()), count * sizeof(T));
}

template <std::size_t TPB, std::size_t IPT>
static void dump_mp(const std::string& outdir, int* off, std::size_t R, std::size_t N) {
  using setup_t = schedule::setup<schedule::algorithms_t::merge_path_flat, TPB, IPT, int,
                                  int, std::size_t, std::size_t>;
  const std::size_t tiles = math::ceil_div(R + N, TPB * IPT);
  const std::string tag = std::to_string(TPB) + "_" + std::to_string(IPT);
  std::cout << "tiles_" << tag << " " << tiles << std::endl;
  if (tiles == 0) return;
  vector_t<unsigned> out(tiles * TPB * 2);
  for (int with_meta = 0; with_meta < 2; ++with_meta) {
    (void)hipMemset(out.raw(), 0xff, out.size() * sizeof(unsigned));
    // preprocess_t computes the coordinates only when its search grid would
    // fill the device (merge_path_flat.hxx:185-205); the in-kernel search
    // stands in otherwise, so the "meta" file is written from whichever the
    // reference's rule picks.
    typename setup_t::meta_t meta(off, R, N);
    if (!with_meta) {
      typename setup_t::meta_t none(off, 0, 0);  // no coordinates: every tile searches
      launch::non_cooperative(0, k_mp_coords<TPB, IPT, typename setup_t::meta_t>,
                              dim3((unsigned)tiles), dim3(TPB), none, R, N, off, out.raw());
    } else {
      launch::non_cooperative(0, k_mp_coords<TPB, IPT, typename setup_t::meta_t>,
                              dim3((unsigned)tiles), dim3(TPB), meta, R, N, off, out.raw());
    }
    error::throw_if_exception(hipDeviceSynchronize(), "k_mp_coords");
    write_bin(outdir + "/mp_" + tag + (with_meta ? "_meta" : "") + ".bin", out.raw(),
              out.size());
    if (with_meta) std::cout << "meta_" << tag << " " << (meta.data() ? 1 : 0) << std::endl;
  }
}

template <std::size_t TPB, std::size_t TPT>
static void dump_gm(const std::string& outdir, int* off, std::size_t R, std::size_t N) {
  vector_t<long long> out(2 * N);
  if (N) (void)hipMemset(out.raw(), 0xff, out.size() * sizeof(long long));
  const unsigned grid = (unsigned)math::ceil_div(R, TPB);
  if (grid)
    launch::non_cooperative(0, k_gm_map<TPB, TPT>, dim3(grid), dim3(TPB), R, N, off,
                            out.raw());
  error::throw_if_exception(hipDeviceSynchronize(), "k_gm_map");
  write_bin(outdir + "/gm_" + std::to_string(TPB) + "_" + std::to_string(TPT) + ".bin",
            out.raw(), out.size());
}

int main(int argc, char** argv) {
  std::string filename, outdir = ".";
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-m" && i + 1 < argc) filename = argv[++i];
    else if (a == "-o" && i + 1 < argc) outdir = argv[++i];
  }
  if (filename.empty()) {
    std::cout << argv[0] << " -m matrix.mtx -o outdir" << std::endl;
    return 0;
  }
  matrix_market_t<int, int, float> mtx;
  csr_t<int, int, float> csr(mtx.load(filename));
  const std::size_t R = csr.rows, N = csr.nnzs;
  int* off = csr.offsets.raw();

  dump_mp<128, 5>(outdir, off, R, N);
  dump_mp<64, 3>(outdir, off, R, N);

  device::properties_t props;
  const unsigned wo_grid = 2 * (unsigned)props.multi_processor_count();
  vector_t<unsigned> wo(std::size_t(wo_grid) * 128 * 4);
  launch::non_cooperative(0, k_wo_coords<128>, dim3(wo_grid), dim3(128), R, N, off, wo.raw());
  error::throw_if_exception(hipDeviceSynchronize(), "k_wo_coords");
  write_bin(outdir + "/wo_128.bin", wo.raw(), wo.size());
  std::cout << "wo_threads " << std::size_t(wo_grid) * 128 << std::endl;

  dump_gm<128, 128>(outdir, off, R, N);
  dump_gm<256, 32>(outdir, off, R, N);
  dump_gm<256, 64>(outdir, off, R, N);
  dump_gm<256, 128>(outdir, off, R, N);
  dump_gm<256, 8>(outdir, off, R, N);
  std::cout << "done" << std::endl;
  return 0;
}
