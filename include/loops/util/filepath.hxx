/**
 * @file filepath.hxx
 * @brief File-name helpers of the example programs (reference
 * include/loops/util/filepath.hxx:14-35): the file name of a path, the
 * dataset name of a file name, and the input kind by suffix.
 *
 * Names too short for a suffix test are "not of that kind" here (the
 * reference's substr on them throws std::out_of_range).
 */
#pragma once

#include <string>

namespace loops {

/// "dir/name.mtx" -> "name.mtx": what follows the last character of delim.
inline std::string extract_filename(const std::string& path,
                                    const std::string& delim = "/") {
  const std::size_t s = path.find_last_of(delim);
  return s == std::string::npos ? path : path.substr(s + 1);
}

/// "name.mtx" -> "name": everything before the last '.'.
inline std::string extract_dataset(const std::string& filename) {
  const std::size_t d = filename.find_last_of('.');
  return d == std::string::npos ? filename : filename.substr(0, d);
}

namespace detail {
inline bool ends_with(const std::string& s, const char* suffix) {
  const std::string t(suffix);
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}
}  // namespace detail

/// Matrix Market input: ".mtx" or ".mmio".
inline bool is_market(const std::string& filename) {
  return detail::ends_with(filename, ".mtx") || detail::ends_with(filename, ".mmio");
}

/// Binary CSR input: ".csr".
inline bool is_binary_csr(const std::string& filename) {
  return detail::ends_with(filename, ".csr");
}

}  // namespace loops
