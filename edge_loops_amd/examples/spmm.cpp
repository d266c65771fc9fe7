// SpMM driver (reference examples/spmm/thread_mapped.cu:17-44):
//   loops.spmm.<schedule> -m matrix.mtx [-n columns]
// C = A B with B cols x n (default 10) uniform in [1, 10); prints the
// elapsed time.  Like the reference it does not validate (the tests do).
#include <cstdint>
#include <iostream>
#include <string>

#include <loops/algorithms/spmm/group_mapped.cuh>
#include <loops/algorithms/spmm/merge_path.cuh>
#include <loops/algorithms/spmm/thread_mapped.cuh>
#include <loops/container/market.hxx>
#include <loops/util/timer.hxx>

#ifndef LOOPS_EXAMPLE_SCHEDULE
#define LOOPS_EXAMPLE_SCHEDULE thread_mapped
#endif

using namespace loops;

int main(int argc, char** argv) {
  std::string filename;
  std::size_t n = 10;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if ((a == "-m" || a == "--market") && i + 1 < argc) filename = argv[++i];
    else if (a == "-n" && i + 1 < argc) n = std::stoul(argv[++i]);
  }
  if (filename.empty()) {
    std::cout << argv[0] << " -m matrix.mtx [-n columns]" << std::endl;
    return 0;
  }
  matrix_market_t<int, int, float> mtx;
  csr_t<int, int, float> csr(mtx.load(filename));
  matrix_t<float> B(csr.cols, n), C(csr.rows, n);
  vector_t<float, memory_space_t::host> bh(csr.cols * n);
  for (std::size_t i = 0; i < bh.size(); ++i) bh[i] = 1.0f + float(i % 9);
  (void)hipMemcpy(B.m_data_ptr, bh.data(), bh.size() * sizeof(float), hipMemcpyHostToDevice);
  util::timer_t timer;
  timer.start();
  algorithms::spmm::LOOPS_EXAMPLE_SCHEDULE(csr, B, C);
  timer.stop();
  std::cout << "Elapsed (ms):\t" << timer.milliseconds() << std::endl;
  return 0;
}
