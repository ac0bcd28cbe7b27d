// Row-major FP64 matrix (base/src/sgpp/base/datatypes/DataMatrix.hpp:28).
#pragma once

#include <cstddef>
#include <vector>

#include "sgpp/base/datatypes/DataVector.hpp"
#include "sgpp/base/exception/exceptions.hpp"

namespace sgpp {
namespace base {

class DataMatrix : public std::vector<double> {
 public:
  DataMatrix() = default;
  DataMatrix(size_t nrows, size_t ncols, double v = 0.0)
      : std::vector<double>(nrows * ncols, v), nrows_(nrows), ncols_(ncols) {}
  DataMatrix(const double* p, size_t nrows, size_t ncols)
      : std::vector<double>(p, p + nrows * ncols), nrows_(nrows), ncols_(ncols) {}

  size_t getNrows() const { return nrows_; }
  size_t getNcols() const { return ncols_; }
  double* getPointer() { return data(); }
  const double* getPointer() const { return data(); }
  double get(size_t r, size_t c) const { return (*this)[r * ncols_ + c]; }
  void set(size_t r, size_t c, double v) { (*this)[r * ncols_ + c] = v; }
  void getRow(size_t r, DataVector& v) const {
    v.resize(ncols_);
    for (size_t c = 0; c < ncols_; ++c) v[c] = get(r, c);
  }
  void setRow(size_t r, const DataVector& v) {
    for (size_t c = 0; c < ncols_; ++c) set(r, c, v[c]);
  }
  void resize(size_t nrows) {
    std::vector<double>::resize(nrows * ncols_);
    nrows_ = nrows;
  }
  void resize(size_t nrows, size_t ncols) {
    std::vector<double>::resize(nrows * ncols);
    nrows_ = nrows;
    ncols_ = ncols;
  }
  void transpose() {
    std::vector<double> t(size());
    for (size_t r = 0; r < nrows_; ++r)
      for (size_t c = 0; c < ncols_; ++c) t[c * nrows_ + r] = (*this)[r * ncols_ + c];
    std::vector<double>::swap(t);
    const size_t n = nrows_;
    nrows_ = ncols_;
    ncols_ = n;
  }

 private:
  size_t nrows_ = 0, ncols_ = 0;
};

}  // namespace base
}  // namespace sgpp
