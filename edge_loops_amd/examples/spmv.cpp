// SpMV command-line driver built once per schedule (loops.spmv.<name>),
// mirroring the reference's examples/spmv/<alg>.cu and helpers.hxx:
//   loops.spmv.merge_path -m matrix.mtx [--validate] [-v] [-t float|double|int]
// prints "<kernel>,<dataset>,rows,cols,nnzs,elapsed_ms" and, with
// --validate, recomputes y on the host in type_t, CSR order, and prints the
// number of entries differing by more than 1e-2 (helpers.hxx:97-146).
// x is uniform in [1, 10) like the reference's example (seeded here; whole
// numbers for int).  -t picks type_t (default float), the value types of the
// reference's benchmark (benchmarks/spmv/work_oriented.cu:53).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

#include <loops/algorithms/spmv/group_mapped.cuh>
#include <loops/algorithms/spmv/merge_path_flat.cuh>
#include <loops/algorithms/spmv/original.cuh>
#include <loops/algorithms/spmv/thread_mapped.cuh>
#include <loops/algorithms/spmv/work_oriented.cuh>
#include <loops/container/market.hxx>
#include <loops/util/equal.hxx>
#include <loops/util/timer.hxx>

#ifndef LOOPS_EXAMPLE_SCHEDULE
#define LOOPS_EXAMPLE_SCHEDULE merge_path_flat
#endif
#define LOOPS_STR2(s) #s
#define LOOPS_STR(s) LOOPS_STR2(s)

using namespace loops;

struct parameters_t {
  std::string filename;
  std::string type = "float";
  bool validate = false;
  bool verbose = false;
  parameters_t(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      if ((a == "-m" || a == "--market") && i + 1 < argc) filename = argv[++i];
      else if ((a == "-t" || a == "--type") && i + 1 < argc) type = argv[++i];
      else if (a == "--validate") validate = true;
      else if (a == "-v" || a == "--verbose") verbose = true;
      else usage(argv[0], 0);
    }
    if (filename.empty() || !is_market(filename)) usage(argv[0], 0);
    if (type != "float" && type != "double" && type != "int") usage(argv[0], 1);
  }
  [[noreturn]] static void usage(const char* prog, int code) {
    std::cout << prog << ": Sparse Matrix-Vector Multiplication\n"
              << "  -h, --help        Print help\n"
              << "  -m, --market arg  Matrix file\n"
              << "      --validate    CPU validation\n"
              << "  -t, --type arg    float (default), double or int values" "\n"
              << "  -v, --verbose     Verbose output" << std::endl;
    std::exit(code);
  }
};

template <typename type_t>
int run(const parameters_t& parameters) {
  using index_t = int;
  using offset_t = int;

  matrix_market_t<index_t, offset_t, type_t> mtx;
  csr_t<index_t, offset_t, type_t> csr(mtx.load(parameters.filename));
  vector_t<type_t> x(csr.cols), y(csr.rows);

  // x uniform in [1, 10), seeded (splitmix64 of the index).
  vector_t<type_t, memory_space_t::host> xh(csr.cols);
  for (std::size_t i = 0; i < csr.cols; ++i) {
    uint64_t z = i + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    const double u = double(z >> 40) / 16777216.0;
    xh[i] = std::is_same<type_t, int>::value ? type_t(1 + int(9.0 * u)) : type_t(1.0 + 9.0 * u);
  }
  x = vector_t<type_t>(xh);

  float elapsed = 0.f;
  {
    util::timer_t t;
    t.start();
#define RUN(name) algorithms::spmv::name(csr, x, y)
    auto sched = std::string(LOOPS_STR(LOOPS_EXAMPLE_SCHEDULE));
    if (sched == "merge_path_flat") {
      elapsed = algorithms::spmv::merge_path_flat(csr, x, y).milliseconds();
    } else {
      RUN(LOOPS_EXAMPLE_SCHEDULE);
      elapsed = t.stop();
    }
  }
  std::cout << LOOPS_STR(LOOPS_EXAMPLE_SCHEDULE) << "," << mtx.dataset << ","
            << csr.rows << "," << csr.cols << "," << csr.nnzs << "," << elapsed
            << std::endl;

  if (parameters.validate) {
    csr_t<index_t, offset_t, type_t, memory_space_t::host> h(csr);
    std::vector<type_t> yh(csr.rows);
    for (std::size_t row = 0; row < csr.rows; ++row) {
      type_t sum = 0;
      for (auto nz = h.offsets[row]; nz < h.offsets[row + 1]; ++nz)
        sum += h.values[nz] * xh[h.indices[nz]];
      yh[row] = sum;
    }
    std::size_t errors = util::equal(
        y.raw(), yh.data(), csr.rows,
        [](const type_t a, const type_t b) { return std::abs(double(a) - double(b)) > 1e-2; },
        parameters.verbose);
    std::cout << "Matrix:\t\t" << extract_filename(parameters.filename) << std::endl;
    std::cout << "Dimensions:\t" << csr.rows << " x " << csr.cols << " ("
              << csr.nnzs << ")" << std::endl;
    std::cout << "Errors:\t\t" << errors << std::endl;
  }
  return 0;
}

int main(int argc, char** argv) {
  parameters_t parameters(argc, argv);
  if (parameters.type == "double") return run<double>(parameters);
  if (parameters.type == "int") return run<int>(parameters);
  return run<float>(parameters);
}
