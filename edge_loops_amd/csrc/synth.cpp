// Seeded synthetic CSR generator for the benchmark and test workloads
// (host C++, libloops_synth.so).  Deterministic for a given seed regardless
// of the number of host threads: every row draws from its own counter-based
// stream (splitmix64 of seed and row id).
//
//   row lengths  truncated discrete Zipf(s) on [1, max_len] (or a fixed
//                length when s == 0), scaled so the total is close to
//                target_nnz (min length min_len), optionally with row 0 set
//                to the maximum length;
//   columns      distinct, uniform on [0, cols), sorted per row;
//   values, x    uniform [-1, 1) with 24-bit resolution.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct stream_t {
  uint64_t state;
  explicit stream_t(uint64_t seed, uint64_t id) : state(mix64(seed ^ mix64(id))) {}
  uint64_t next() {
    state += 0x9e3779b97f4a7c15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double unit() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  // Uniform integer in [0, n).
  uint64_t below(uint64_t n) {
    return (uint64_t)(((unsigned __int128)next() * n) >> 64);
  }
  float sym() {  // uniform [-1, 1), 24-bit
    return float(next() >> 40) * (2.0f / 16777216.0f) - 1.0f;
  }
};

template <typename F>
void parallel_rows(int64_t n, int nthreads, F&& f) {
  if (nthreads < 1) nthreads = 1;
  if (n < 4096) nthreads = 1;
  if (nthreads == 1) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    int64_t a = n * t / nthreads, b = n * (t + 1) / nthreads;
    th.emplace_back([=, &f] { f(a, b); });
  }
  for (auto& t : th) t.join();
}

}  // namespace

extern "C" {

// Row lengths.  Returns 0 and writes lengths[rows] and *nnz_out, or -1 on
// invalid arguments / int32 overflow of the total.
int loops_synth_row_lengths(int64_t rows, int64_t cols, double zipf_s,
                            int64_t min_len, int64_t max_len, int64_t target_nnz,
                            int force_first_max, uint64_t seed, int32_t* lengths,
                            int64_t* nnz_out, int nthreads) {
  if (rows < 0 || cols < 0 || min_len < 0 || max_len < min_len || !lengths)
    return -1;
  int64_t cap = std::min<int64_t>(max_len, cols);
  int64_t lo = std::min<int64_t>(min_len, cap);
  std::vector<double> raw(rows);
  if (zipf_s <= 0.0) {
    for (int64_t r = 0; r < rows; ++r) raw[r] = double(cap);
  } else {
    // CDF of P(k) ~ k^-s on k = 1..max_len.
    std::vector<double> cdf(max_len);
    double acc = 0;
    for (int64_t k = 1; k <= max_len; ++k) {
      acc += std::pow(double(k), -zipf_s);
      cdf[k - 1] = acc;
    }
    for (auto& c : cdf) c /= acc;
    parallel_rows(rows, nthreads, [&](int64_t a, int64_t b) {
      for (int64_t r = a; r < b; ++r) {
        stream_t st(seed ^ 0x5a5a5a5aull, uint64_t(r));
        double u = st.unit();
        int64_t k = std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin() + 1;
        raw[r] = double(std::min<int64_t>(k, max_len));
      }
    });
  }
  auto realize = [&](double scale, int64_t r) -> int64_t {
    if (force_first_max && r == 0) return cap;
    int64_t L = (int64_t)std::llround(raw[r] * scale);
    return std::max<int64_t>(lo, std::min<int64_t>(cap, L));
  };
  auto total_for = [&](double scale) {
    int64_t t = 0;
    for (int64_t r = 0; r < rows; ++r) t += realize(scale, r);
    return t;
  };
  double scale = 1.0;
  if (target_nnz > 0 && zipf_s > 0.0 && rows > 0) {
    double a = 0.0, b = 1.0;
    while (total_for(b) < target_nnz && b < 1e6) b *= 2.0;
    for (int it = 0; it < 60; ++it) {
      double m = 0.5 * (a + b);
      if (total_for(m) < target_nnz) a = m;
      else b = m;
    }
    scale = b;
  }
  int64_t t = 0;
  for (int64_t r = 0; r < rows; ++r) {
    int64_t L = realize(scale, r);
    lengths[r] = (int32_t)L;
    t += L;
  }
  if (t + rows >= INT32_MAX) return -1;
  *nnz_out = t;
  return 0;
}

// Fills distinct sorted column ids and values for every row, given offsets
// (rows + 1 entries, prefix sums of the lengths).
int loops_synth_fill(int64_t rows, int64_t cols, const int32_t* offsets,
                     uint64_t seed, int32_t* indices, float* values,
                     int nthreads) {
  if (rows < 0 || cols < 0 || !offsets) return -1;
  for (int64_t r = 0; r < rows; ++r)
    if (offsets[r + 1] - offsets[r] > cols || offsets[r + 1] < offsets[r]) return -1;
  parallel_rows(rows, nthreads, [&](int64_t a, int64_t b) {
    std::vector<int32_t> buf;
    for (int64_t r = a; r < b; ++r) {
      const int64_t L = offsets[r + 1] - offsets[r];
      int32_t* out = indices + offsets[r];
      stream_t st(seed, uint64_t(r));
      if (L == 0) continue;
      if (4 * L >= cols) {
        // Selection sampling over all columns (already sorted).
        int64_t need = L, left = cols, w = 0;
        for (int64_t c = 0; c < cols && need > 0; ++c, --left) {
          if ((int64_t)st.below((uint64_t)left) < need) {
            out[w++] = (int32_t)c;
            --need;
          }
        }
      } else {
        buf.clear();
        while ((int64_t)buf.size() < L) {
          int64_t missing = L - (int64_t)buf.size();
          for (int64_t i = 0; i < missing; ++i)
            buf.push_back((int32_t)st.below((uint64_t)cols));
          std::sort(buf.begin(), buf.end());
          buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        }
        std::copy(buf.begin(), buf.end(), out);
      }
      float* v = values + offsets[r];
      for (int64_t i = 0; i < L; ++i) v[i] = st.sym();
    }
  });
  return 0;
}

// out[i] = uniform [-1, 1) from a counter-based stream.
void loops_synth_uniform(uint64_t seed, int64_t n, float* out, int nthreads) {
  parallel_rows(n, nthreads, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      uint64_t z = mix64(seed ^ mix64(uint64_t(i) + 0x1234567ull));
      out[i] = float(z >> 40) * (2.0f / 16777216.0f) - 1.0f;
    }
  });
}

}  // extern "C"
