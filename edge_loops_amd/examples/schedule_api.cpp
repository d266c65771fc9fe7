// User kernels written against the device-side schedule API
// (include/loops/schedule.hxx), in the style of the reference's
// algorithms/spmv/*.cuh and docs/loadbalancing_api.md, compiled with hipcc
// for gfx950.  Usage:
//   loops.schedule_api -m matrix.mtx -o outdir
// writes outdir/x.bin and outdir/y_<kernel>.bin (raw fp32) for:
//   thread_mapped   setup<thread_mapped>: tiles() / atoms(row)
//   block_mapped    setup<group_mapped, 128, 128>: partition / atom_accessor /
//                   tile_accessor / tile_id / atom_id, atomicAdd per nonzero
//   warp_mapped     the same with 32-lane groups in 256-lane blocks (the
//                   reference's warp_mapped, group_mapped.hxx:239-241)
//   wave_mapped     64-lane groups (one wavefront) in 256-lane blocks
//   group128        128-lane groups (two wavefronts) in 256-lane blocks
//   work_oriented   setup<work_oriented>: init / tiles / atoms / remainder_*,
//                   every partial row added with atomicAdd (the reference
//                   stores complete rows, which races with another block's
//                   atomicAdd of the same row: work_oriented.cuh:60 vs :76)
//   merge_path_flat setup<merge_path_flat, 128, 5> with the preprocess_t
//                   tile coordinates, atomicAdd per nonzero (:38-83)
// The atomic kernels add into y, which starts at zero.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include <loops/container/csr.hxx>
#include <loops/container/market.hxx>
#include <loops/schedule.hxx>
#include <loops/util/device.hxx>
#include <loops/util/launch.hxx>
#include <loops/util/math.hxx>

using namespace loops;

template <typename setup_t>
__global__ void k_thread_mapped(setup_t config, const int* indices,
                                const float* values, const float* x, float* y) {
  for (auto row : config.tiles()) {
    float sum = 0;
    for (auto nz : config.atoms(row)) sum += values[nz] * x[indices[nz]];
    y[row] = sum;
  }
}

template <std::size_t TPB, std::size_t TPT>
__global__ void __launch_bounds__(TPB) k_group_mapped(std::size_t rows, std::size_t nnz,
                                                      int* offsets, const int* indices,
                                                      const float* values,
                                                      const float* x, float* y) {
  using setup_t = schedule::setup<schedule::algorithms_t::group_mapped, TPB, TPT, int, int>;
  __shared__ typename setup_t::storage_t temporary_storage;
  setup_t config(temporary_storage, offsets, rows, nnz);
  auto p = config.partition();
  for (auto virtual_atom : config.atom_accessor(p)) {
    auto virtual_tile = config.tile_accessor(virtual_atom, p);
    if (!(config.is_valid_accessor(virtual_tile, p))) continue;
    auto row = config.tile_id(virtual_tile, p);
    auto nz_idx = config.atom_id(virtual_atom, row, virtual_tile, p);
    atomicAdd(&(y[row]), values[nz_idx] * x[indices[nz_idx]]);
  }
}

template <std::size_t TPB>
__global__ void __launch_bounds__(TPB, 2) k_work_oriented(std::size_t rows, std::size_t nnz,
                                                          int* offsets, const int* indices,
                                                          const float* values,
                                                          const float* x, float* y) {
  using setup_t = schedule::setup<schedule::algorithms_t::work_oriented, TPB, 1, int, int,
                                  std::size_t, std::size_t>;
  setup_t config(offsets, rows, nnz);
  auto map = config.init();
  for (auto row : config.tiles(map)) {
    float sum = 0;
    for (auto nz : config.atoms(row, map)) sum += values[nz] * x[indices[nz]];
    atomicAdd(&(y[row]), sum);
  }
  for (auto row : config.remainder_tiles(map)) {
    float sum = 0;
    for (auto nz : config.remainder_atoms(map)) sum += values[nz] * x[indices[nz]];
    if (sum != 0) atomicAdd(&(y[row]), sum);
  }
}

template <std::size_t TPB, std::size_t IPT, typename meta_t>
__global__ void __launch_bounds__(TPB) k_merge_path_flat(meta_t meta, std::size_t rows,
                                                         std::size_t nnz, int* offsets,
                                                         const int* indices,
                                                         const float* values,
                                                         const float* x, float* y) {
  using setup_t = schedule::setup<schedule::algorithms_t::merge_path_flat, TPB, IPT, int,
                                  int, std::size_t, std::size_t>;
  __shared__ typename setup_t::storage_t temporary_storage;
  setup_t config(meta, temporary_storage, offsets, rows, nnz);
  auto map = config.init();
  if (!config.is_valid_accessor(map)) return;
#pragma unroll
  for (auto item : config.virtual_idx()) {
    auto nz = config.atom_idx(item, map);
    auto row = config.tile_idx(map);
    float nonzero = values[nz] * x[indices[nz]];
    if (config.atoms_counting_it[map.y] < temporary_storage.tile_end_offset[map.x]) {
      atomicAdd(&(y[row]), nonzero);
      map.y++;
    } else {
      map.x++;
    }
  }
}

static void write_bin(const std::string& path, const std::vector<float>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(float));
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
  const std::size_t R = csr.rows, C = csr.cols, N = csr.nnzs;
  std::vector<float> xh(C);
  for (std::size_t i = 0; i < C; ++i) xh[i] = float(int((i * 7919u) % 17u) - 8) / 8.0f;
  vector_t<float> x(C), y(R);
  (void)hipMemcpy(x.raw(), xh.data(), C * sizeof(float), hipMemcpyHostToDevice);
  write_bin(outdir + "/x.bin", xh);
  int* off = csr.offsets.raw();
  const int* idx = csr.indices.raw();
  const float* val = csr.values.raw();
  auto save = [&](const char* name) {
    (void)hipDeviceSynchronize();
    std::vector<float> yh(R);
    (void)hipMemcpy(yh.data(), y.raw(), R * sizeof(float), hipMemcpyDeviceToHost);
    write_bin(outdir + "/y_" + name + ".bin", yh);
    (void)hipMemset(y.raw(), 0, R * sizeof(float));
    std::cout << name << " done" << std::endl;
  };
  const unsigned grid128 = (unsigned)math::ceil_div(R, std::size_t(128));
  {
    using setup_t = schedule::setup<schedule::algorithms_t::thread_mapped, 1, 1, int, int>;
    setup_t config(off, R, N);
    launch::non_cooperative(0, k_thread_mapped<setup_t>, dim3(grid128), dim3(128), config,
                            idx, val, x.raw(), y.raw());
    save("thread_mapped");
  }
  launch::non_cooperative(0, k_group_mapped<128, 128>, dim3(grid128), dim3(128), R, N, off,
                          idx, val, x.raw(), y.raw());
  save("block_mapped");
  const unsigned grid256 = (unsigned)math::ceil_div(R, std::size_t(256));
  launch::non_cooperative(0, k_group_mapped<256, 32>, dim3(grid256), dim3(256), R, N, off,
                          idx, val, x.raw(), y.raw());
  save("warp_mapped");
  launch::non_cooperative(0, k_group_mapped<256, 64>, dim3(grid256), dim3(256), R, N, off,
                          idx, val, x.raw(), y.raw());
  save("wave_mapped");
  launch::non_cooperative(0, k_group_mapped<256, 128>, dim3(grid256), dim3(256), R, N, off,
                          idx, val, x.raw(), y.raw());
  save("group128");
  {
    device::properties_t props;
    launch::non_cooperative(0, k_work_oriented<128>,
                            dim3(2 * props.multi_processor_count()), dim3(128), R, N, off,
                            idx, val, x.raw(), y.raw());
    save("work_oriented");
  }
  if (N > 0) {
    using setup_t = schedule::setup<schedule::algorithms_t::merge_path_flat, 128, 5, int,
                                    int, std::size_t, std::size_t>;
    typename setup_t::meta_t meta(off, R, N);
    const unsigned tiles = (unsigned)math::ceil_div(R + N, std::size_t(640));
    launch::non_cooperative(0, k_merge_path_flat<128, 5, typename setup_t::meta_t>,
                            dim3(tiles), dim3(128), meta, R, N, off, idx, val, x.raw(),
                            y.raw());
    save("merge_path_flat");
  }
  return 0;
}
