/**
 * @file matrix.cuh
 * @brief Dense row-major matrix (reference include/loops/container/
 * matrix.cuh:8-51).  The copy constructor is shallow (shares m_data_ptr) so
 * the object can be passed to kernels by value; element (r, c) lives at
 * m_data_ptr[cols * r + c].
 */
#pragma once

#include <cstddef>

#include <loops/container/vector.hxx>

namespace loops {

template <typename value_t, memory_space_t space = memory_space_t::device>
struct matrix_t {
  std::size_t rows;
  std::size_t cols;
  vector_t<value_t, space> m_data;
  value_t* m_data_ptr;

  matrix_t() : rows(0), cols(0), m_data(), m_data_ptr(nullptr) {}
  matrix_t(std::size_t r, std::size_t c)
      : rows(r), cols(c), m_data(r * c), m_data_ptr(m_data.raw()) {}
  __host__ __device__ matrix_t(const matrix_t& other)
      : rows(other.rows), cols(other.cols), m_data_ptr(other.m_data_ptr) {}

  __host__ __device__ __forceinline__ value_t operator()(int r, int c) const {
    return m_data_ptr[cols * std::size_t(r) + c];
  }
  __host__ __device__ __forceinline__ value_t& operator()(int r, int c) {
    return m_data_ptr[cols * std::size_t(r) + c];
  }
  __host__ __device__ __forceinline__ value_t operator[](std::size_t i) const {
    return m_data_ptr[i];
  }
  __host__ __device__ __forceinline__ value_t& operator[](std::size_t i) {
    return m_data_ptr[i];
  }
};

}  // namespace loops
