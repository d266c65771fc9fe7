/**
 * @file error.hxx
 * @brief Exceptions for host code (reference include/loops/error.hxx:15-45):
 * exception_t and throw_if_exception(hipError_t | bool, message).  The C ABI
 * returns status codes; the C++ layer turns them into these exceptions.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <exception>
#include <string>

namespace loops {
namespace error {

typedef hipError_t error_t;

struct exception_t : std::exception {
  std::string report;
  exception_t(error_t status, std::string message = "") {
    report = hipGetErrorString(status) + std::string("\t: ") + message;
  }
  exception_t(std::string message = "") { report = message; }
  virtual const char* what() const noexcept { return report.c_str(); }
};

inline void throw_if_exception(error_t status, std::string message = "") {
  if (status != hipSuccess) throw exception_t(status, message);
}

inline void throw_if_exception(bool is_exception, std::string message = "") {
  if (is_exception) throw exception_t(message);
}

}  // namespace error
}  // namespace loops
