// Checks of the container utilities of the C++ drop-in (include/loops/
// container/{coo,csc,csr}.hxx, container/detail/convert.hxx, util/
// {sample,generate,filepath}.hxx) against known answers and closed forms.
//
//   loops.containers            host containers only (no GPU needed)
//   loops.containers --device   the same plus device containers
//
// Prints one line per group and "containers: N checks passed"; exits 1 at
// the first failed check, naming it.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <loops/container/container.hxx>
#include <loops/util/filepath.hxx>
#include <loops/util/generate.hxx>
#include <loops/util/sample.hxx>

using namespace loops;

static int checks = 0;
#define CHECK(cond)                                                                  \
  do {                                                                               \
    ++checks;                                                                        \
    if (!(cond)) {                                                                   \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                  \
      std::exit(1);                                                                  \
    }                                                                                \
  } while (0)

template <typename T, memory_space_t S>
static std::vector<T> host_copy(const vector_t<T, S>& v) {
  const vector_t<T, memory_space_t::host> h(v);
  return std::vector<T>(h.raw(), h.raw() + h.size());
}

// The sample matrix: its arrays, and the conversions on them.
template <memory_space_t S>
static void sample_and_convert() {
  auto A = sample::csr<S>();
  CHECK(A.rows == 4 && A.cols == 4 && A.nnzs == 4);
  CHECK((host_copy(A.offsets) == std::vector<int>{0, 0, 2, 3, 4}));
  CHECK((host_copy(A.indices) == std::vector<int>{0, 1, 2, 1}));
  CHECK((host_copy(A.values) == std::vector<float>{5, 8, 3, 6}));
  // y = A 1 = {0, 13, 3, 6}
  const auto off = host_copy(A.offsets);
  const auto val = host_copy(A.values);
  std::vector<float> y(4, 0.f);
  for (int r = 0; r < 4; ++r)
    for (int k = off[r]; k < off[r + 1]; ++k) y[r] += val[k];
  CHECK((y == std::vector<float>{0, 13, 3, 6}));

  // offsets -> row of every entry -> offsets
  vector_t<int, S> rows_of(4);
  detail::offsets_to_indices(A.offsets, rows_of);
  CHECK((host_copy(rows_of) == std::vector<int>{1, 1, 2, 3}));
  vector_t<int, S> back(5);
  detail::indices_to_offsets(rows_of, back);
  CHECK((host_copy(back) == std::vector<int>{0, 0, 2, 3, 4}));

  // COO from CSR, CSC from that COO: (col, row) order
  coo_t<int, float, memory_space_t::host> coo(A);
  CHECK(coo.nnzs == 4);
  CHECK((host_copy(coo.row_indices) == std::vector<int>{1, 1, 2, 3}));
  CHECK((host_copy(coo.col_indices) == std::vector<int>{0, 1, 2, 1}));
  csc_t<int, int, float, S> B(coo);
  CHECK(B.rows == 4 && B.cols == 4 && B.nnzs == 4);
  CHECK((host_copy(B.offsets) == std::vector<int>{0, 1, 3, 4, 4}));
  CHECK((host_copy(B.indices) == std::vector<int>{1, 1, 3, 2}));
  CHECK((host_copy(B.values) == std::vector<float>{5, 8, 6, 3}));
  // and back to CSR
  csr_t<int, int, float, S> C(coo);
  CHECK(host_copy(C.offsets) == off && host_copy(C.indices) == host_copy(A.indices) &&
        host_copy(C.values) == val);
  csc_t<int, int, float, S> E(4, 6, 3);
  CHECK(E.offsets.size() == 7 && E.indices.size() == 3);
  std::printf("sample, conversions, COO / CSC / CSR: ok\n");
}

// Edge cases of the conversions on host arrays.
static void convert_edges() {
  // empty rows at both ends and in the middle; a last row holding the tail
  const int off[] = {0, 0, 3, 3, 3, 5, 5};
  int idx[5];
  detail::offsets_to_indices(off, 7, idx, 5);
  CHECK(idx[0] == 1 && idx[1] == 1 && idx[2] == 1 && idx[3] == 4 && idx[4] == 4);
  int off2[7];
  detail::indices_to_offsets(idx, 5, off2, 7);
  CHECK(std::memcmp(off, off2, sizeof(off)) == 0);
  // no entries at all
  const int zero[] = {0, 0, 0};
  int none[1] = {7};
  detail::offsets_to_indices(zero, 3, none, 0);
  CHECK(none[0] == 7);
  int off3[3] = {9, 9, 9};
  detail::indices_to_offsets(none, 0, off3, 3);
  CHECK(off3[0] == 0 && off3[1] == 0 && off3[2] == 0);
  std::printf("conversion edge cases: ok\n");
}

// Sorts are stable; remove_duplicates keeps the first entry of a pair.
static void coo_sorts() {
  coo_t<int, float, memory_space_t::host> c(3, 4, 6);
  const int r[] = {2, 0, 2, 1, 0, 2};
  const int k[] = {3, 1, 3, 0, 1, 0};
  const float v[] = {10, 20, 30, 40, 50, 60};
  for (int i = 0; i < 6; ++i) {
    c.row_indices[i] = r[i];
    c.col_indices[i] = k[i];
    c.values[i] = v[i];
  }
  auto byrow = c;
  byrow.sort_by_row();
  CHECK((host_copy(byrow.row_indices) == std::vector<int>{0, 0, 1, 2, 2, 2}));
  CHECK((host_copy(byrow.col_indices) == std::vector<int>{1, 1, 0, 0, 3, 3}));
  CHECK((host_copy(byrow.values) == std::vector<float>{20, 50, 40, 60, 10, 30}));
  auto bycol = c;
  bycol.sort_by_column();
  CHECK((host_copy(bycol.col_indices) == std::vector<int>{0, 0, 1, 1, 3, 3}));
  CHECK((host_copy(bycol.row_indices) == std::vector<int>{1, 2, 0, 0, 2, 2}));
  CHECK((host_copy(bycol.values) == std::vector<float>{40, 60, 20, 50, 10, 30}));
  auto u = c;
  u.remove_duplicates();
  CHECK(u.nnzs == 4 && u.values.size() == 4);
  CHECK((host_copy(u.row_indices) == std::vector<int>{0, 1, 2, 2}));
  CHECK((host_copy(u.col_indices) == std::vector<int>{1, 0, 0, 3}));
  CHECK((host_copy(u.values) == std::vector<float>{20, 40, 60, 10}));
  std::printf("COO sorts and duplicate removal: ok\n");
}

static void file_names() {
  CHECK(extract_filename("datasets/chesapeake/chesapeake.mtx") == "chesapeake.mtx");
  CHECK(extract_filename("chesapeake.mtx") == "chesapeake.mtx");
  CHECK(extract_dataset("chesapeake.mtx") == "chesapeake");
  CHECK(extract_dataset("a.b.mtx") == "a.b");
  CHECK(is_market("x.mtx") && is_market("x.mmio") && !is_market("x.csr") && !is_market("tx"));
  CHECK(is_binary_csr("graph.csr") && !is_binary_csr("graph.mtx") && !is_binary_csr("sr"));
  std::printf("file names: ok\n");
}

// Fills: bounds, both integer ends reached, reproducible by seed, different
// across seeds, the mean of a [-1, 1) fill near 0; host and device equal.
template <memory_space_t S>
static void fills() {
  const std::size_t n = 1 << 16;
  vector_t<float, S> f(n);
  generate::random::uniform_distribution(f.begin(), f.end(), -1.f, 1.f, 1234u);
  const auto a = host_copy(f);
  double mean = 0;
  for (float x : a) {
    CHECK(x >= -1.f && x < 1.f);
    mean += x;
  }
  mean /= n;
  CHECK(mean > -0.02 && mean < 0.02);
  vector_t<float, S> g(n);
  generate::random::uniform_distribution(g.begin(), g.end(), -1.f, 1.f, 1234u);
  CHECK(host_copy(g) == a);
  generate::random::uniform_distribution(g.begin(), g.end(), -1.f, 1.f, 1235u);
  std::size_t same = 0;
  const auto b = host_copy(g);
  for (std::size_t i = 0; i < n; ++i) same += a[i] == b[i];
  CHECK(same < n / 100);
  vector_t<int, S> q(n);
  generate::random::uniform_distribution(q.begin(), q.end(), 3, 7, 99u);
  int lo = 100, hi = -100;
  std::size_t count[5] = {0, 0, 0, 0, 0};
  for (int x : host_copy(q)) {
    CHECK(x >= 3 && x <= 7);
    lo = x < lo ? x : lo;
    hi = x > hi ? x : hi;
    ++count[x - 3];
  }
  CHECK(lo == 3 && hi == 7);
  for (std::size_t c : count) CHECK(c > n / 5 - n / 50 && c < n / 5 + n / 50);
  // the same fill on the host
  std::vector<float> h(n);
  generate::random::uniform_distribution(h.data(), h.data() + n, -1.f, 1.f, 1234u);
  CHECK(h == a);
  std::printf("uniform fills (%s): ok\n", S == memory_space_t::host ? "host" : "device");
}

// Random CSR matrices: well formed, no repeated pair, in range.
template <memory_space_t S>
static void random_csr() {
  csr_t<int, int, float, S> m;
  generate::random::csr(300, 200, 0.05f, m, 7u);
  CHECK(m.rows == 300 && m.cols == 200);
  CHECK(m.nnzs > 2500 && m.nnzs <= 3000);  // 3000 drawn, about 3 % repeats
  const auto off = host_copy(m.offsets);
  const auto idx = host_copy(m.indices);
  const auto val = host_copy(m.values);
  CHECK(off.size() == 301 && off[0] == 0 && (std::size_t)off[300] == m.nnzs);
  for (int r = 0; r < 300; ++r) {
    CHECK(off[r] <= off[r + 1]);
    for (int k = off[r]; k < off[r + 1]; ++k) {
      CHECK(idx[k] >= 0 && idx[k] < 200);
      if (k > off[r]) CHECK(idx[k - 1] < idx[k]);
      CHECK(val[k] >= 0.f && val[k] < 1.f);
    }
  }
  csr_t<int, int, float, S> m2;
  generate::random::csr(300, 200, 0.05f, m2, 7u);
  CHECK(host_copy(m2.indices) == idx && host_copy(m2.values) == val);
  csr_t<int, int, float, S> empty;
  generate::random::csr(0, 10, 0.5f, empty, 7u);
  CHECK(empty.rows == 0 && empty.nnzs == 0 && empty.offsets.size() == 1);
  std::printf("random CSR (%s): ok\n", S == memory_space_t::host ? "host" : "device");
}

int main(int argc, char** argv) {
  const bool device = argc > 1 && std::strcmp(argv[1], "--device") == 0;
  sample_and_convert<memory_space_t::host>();
  convert_edges();
  coo_sorts();
  file_names();
  fills<memory_space_t::host>();
  random_csr<memory_space_t::host>();
  if (device) {
    sample_and_convert<memory_space_t::device>();
    fills<memory_space_t::device>();
    random_csr<memory_space_t::device>();
  }
  std::printf("containers: %d checks passed\n", checks);
  return 0;
}
