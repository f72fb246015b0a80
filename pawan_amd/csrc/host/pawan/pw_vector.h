// pw_vector.h -- the plain double containers the class API uses where the
// reference uses gsl_vector / gsl_matrix (src/wake/wake.h:49-57): a vector
// of `size` doubles and a row-major size1 x size2 matrix (tda == size2),
// both owning contiguous storage.
#pragma once

#include <cstddef>
#include <cstring>
#include <vector>

namespace pawan {

struct vec {
  size_t size = 0;
  std::vector<double> data;
  vec() = default;
  explicit vec(size_t n) : size(n), data(n, 0.0) {}
  double get(size_t i) const { return data[i]; }
  void set(size_t i, double v) { data[i] = v; }
  void set_zero() { std::fill(data.begin(), data.end(), 0.0); }
  double* ptr() { return data.data(); }
  const double* ptr() const { return data.data(); }
};

struct mat {
  size_t size1 = 0, size2 = 0;
  std::vector<double> data;
  mat() = default;
  mat(size_t rows, size_t cols) : size1(rows), size2(cols), data(rows * cols, 0.0) {}
  double get(size_t i, size_t j) const { return data[i * size2 + j]; }
  void set(size_t i, size_t j, double v) { data[i * size2 + j] = v; }
  void set_zero() { std::fill(data.begin(), data.end(), 0.0); }
  double* row(size_t i) { return data.data() + i * size2; }
  const double* row(size_t i) const { return data.data() + i * size2; }
  double* ptr() { return data.data(); }
  const double* ptr() const { return data.data(); }
};

}  // namespace pawan
