// FP64 vector with the member functions the operator boundary and the CG
// solver use (base/src/sgpp/base/datatypes/DataVector.hpp:25).
#pragma once

#include <cstddef>
#include <vector>

namespace sgpp {
namespace base {

class DataVector : public std::vector<double> {
 public:
  DataVector() = default;
  explicit DataVector(size_t n, double v = 0.0) : std::vector<double>(n, v) {}
  DataVector(const double* p, size_t n) : std::vector<double>(p, p + n) {}

  size_t getSize() const { return size(); }
  double* getPointer() { return data(); }
  const double* getPointer() const { return data(); }
  double get(size_t i) const { return (*this)[i]; }
  void set(size_t i, double v) { (*this)[i] = v; }
  void setAll(double v) {
    for (double& x : *this) x = v;
  }
  void copyFrom(const DataVector& o) {
    // SG++ copies min(size) elements (DataVector.cpp copyFrom)
    const size_t n = o.size() < size() ? o.size() : size();
    for (size_t i = 0; i < n; ++i) (*this)[i] = o[i];
  }
  // serial dot product, DataVector.cpp:249-257
  double dotProduct(const DataVector& o) const {
    double s = 0.0;
    for (size_t i = 0; i < size(); ++i) s += (*this)[i] * o[i];
    return s;
  }
  // this += a * x; silently ignores size mismatch as SG++ does (DataVector.cpp:370-378)
  void axpy(double a, const DataVector& x) {
    if (x.size() != size()) return;
    for (size_t i = 0; i < size(); ++i) (*this)[i] += a * x[i];
  }
  void sub(const DataVector& x) {
    for (size_t i = 0; i < size() && i < x.size(); ++i) (*this)[i] -= x[i];
  }
  void add(const DataVector& x) {
    for (size_t i = 0; i < size() && i < x.size(); ++i) (*this)[i] += x[i];
  }
  void mult(double s) {
    for (double& v : *this) v *= s;
  }
};

}  // namespace base
}  // namespace sgpp
