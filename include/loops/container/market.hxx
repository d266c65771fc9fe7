/**
 * @file market.hxx
 * @brief Matrix Market (.mtx) reader (reference include/loops/container/
 * market.hxx:99-227, detail/mmio.cpp:101-213).
 *
 * Reads the banner ("%%MatrixMarket matrix coordinate <pattern|real|integer>
 * <general|symmetric>"), the size line and the entries; converts 1-based to
 * 0-based ids; pattern entries get value 1; symmetric files get every
 * off-diagonal entry mirrored right after it, in file order.  Errors throw
 * error::exception_t (the reference exits the process for unreadable files
 * and bad banners, and throws for zero-based ids and size overflow).
 */
#pragma once

#include <cctype>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>

#include <loops/container/coo.hxx>
#include <loops/error.hxx>
#include <loops/util/filepath.hxx>

namespace loops {

enum matrix_market_format_t { coordinate, array };
enum matrix_market_data_t { real, complex, pattern, integer };
enum matrix_market_storage_scheme_t { general, hermitian, symmetric, skew_symmetric };

template <typename index_t, typename offset_t, typename type_t>
struct matrix_market_t {
  std::string filename;
  std::string dataset;
  matrix_market_format_t format = coordinate;
  matrix_market_data_t data = real;
  matrix_market_storage_scheme_t scheme = general;

  coo_t<index_t, type_t, memory_space_t::host> load(std::string _filename) {
    filename = _filename;
    dataset = extract_dataset(extract_filename(filename));
    FILE* file = std::fopen(filename.c_str(), "r");
    error::throw_if_exception(file == nullptr,
                              "File could not be opened: " + filename);
    struct closer {
      FILE* f;
      ~closer() { std::fclose(f); }
    } guard{file};

    char line[4096];
    error::throw_if_exception(std::fgets(line, sizeof line, file) == nullptr,
                              "Could not process Matrix Market banner");
    std::string banner(line);
    for (auto& ch : banner) ch = char(std::tolower((unsigned char)ch));
    error::throw_if_exception(banner.rfind("%%matrixmarket", 0) != 0 ||
                                  banner.find("matrix") == std::string::npos,
                              "Could not process Matrix Market banner");
    format = banner.find("coordinate") != std::string::npos ? coordinate : array;
    error::throw_if_exception(format != coordinate,
                              "Unrecognized matrix market format type");
    if (banner.find("pattern") != std::string::npos) data = pattern;
    else if (banner.find("integer") != std::string::npos) data = integer;
    else if (banner.find("real") != std::string::npos) data = real;
    else error::throw_if_exception(true, "Unrecognized matrix market format type");
    scheme = banner.find("skew-symmetric") != std::string::npos ? skew_symmetric
             : banner.find("symmetric") != std::string::npos   ? symmetric
             : banner.find("hermitian") != std::string::npos   ? hermitian
                                                               : general;

    do {
      error::throw_if_exception(std::fgets(line, sizeof line, file) == nullptr,
                                "Could not read file info (M, N, NNZ)");
    } while (line[0] == '%');
    unsigned long long num_rows = 0, num_columns = 0, num_nonzeros = 0;
    error::throw_if_exception(
        std::sscanf(line, "%llu %llu %llu", &num_rows, &num_columns,
                    &num_nonzeros) != 3,
        "Could not read file info (M, N, NNZ)");
    error::throw_if_exception(
        num_rows >= (unsigned long long)std::numeric_limits<index_t>::max() ||
            num_columns >= (unsigned long long)std::numeric_limits<index_t>::max(),
        "index_t overflow");
    error::throw_if_exception(
        num_nonzeros >= (unsigned long long)std::numeric_limits<offset_t>::max(),
        "offset_t overflow");

    coo_t<index_t, type_t, memory_space_t::host> coo(num_rows, num_columns,
                                                     num_nonzeros);
    for (std::size_t i = 0; i < num_nonzeros; ++i) {
      unsigned long long r = 0, c = 0;
      double w = 1.0;
      int got = data == pattern ? std::fscanf(file, " %llu %llu", &r, &c)
                                : std::fscanf(file, " %llu %llu %lf", &r, &c, &w);
      error::throw_if_exception(got != (data == pattern ? 2 : 3),
                                "Could not read edge from market file");
      error::throw_if_exception(r == 0 || c == 0, "Market file is zero-indexed");
      error::throw_if_exception(r > num_rows || c > num_columns,
                                "Market entry out of range");
      coo.row_indices[i] = index_t(r - 1);
      coo.col_indices[i] = index_t(c - 1);
      coo.values[i] = type_t(w);
    }

    if (scheme == symmetric) {
      std::size_t off = 0;
      for (std::size_t i = 0; i < coo.nnzs; ++i)
        off += coo.row_indices[i] != coo.col_indices[i];
      std::size_t n = coo.nnzs + off;
      error::throw_if_exception(
          n >= (std::size_t)std::numeric_limits<offset_t>::max(),
          "offset_t overflow");
      vector_t<index_t, memory_space_t::host> I(n), J(n);
      vector_t<type_t, memory_space_t::host> V(n);
      std::size_t p = 0;
      for (std::size_t i = 0; i < coo.nnzs; ++i) {
        I[p] = coo.row_indices[i];
        J[p] = coo.col_indices[i];
        V[p] = coo.values[i];
        ++p;
        if (coo.row_indices[i] != coo.col_indices[i]) {
          I[p] = coo.col_indices[i];
          J[p] = coo.row_indices[i];
          V[p] = coo.values[i];
          ++p;
        }
      }
      coo.row_indices = std::move(I);
      coo.col_indices = std::move(J);
      coo.values = std::move(V);
      coo.nnzs = n;
    }
    return coo;
  }
};

}  // namespace loops
