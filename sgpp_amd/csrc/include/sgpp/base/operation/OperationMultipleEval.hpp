// Abstract multi-evaluation operator, same members and contract as
// base/src/sgpp/base/operation/hash/OperationMultipleEval.hpp:27-130.
// B in R^{M x N}, B_ij = phi_j(x_i): mult computes B*alpha (M values),
// multTranspose computes B^T*source (N values).
#pragma once

#include <cstddef>
#include <string>

#include "sgpp/base/datatypes/DataMatrix.hpp"
#include "sgpp/base/datatypes/DataVector.hpp"
#include "sgpp/base/exception/exceptions.hpp"
#include "sgpp/base/grid/Grid.hpp"

namespace sgpp {
namespace base {

class OperationMultipleEval {
 protected:
  Grid& grid;
  DataMatrix& dataset;
  bool isPrepared;

 public:
  OperationMultipleEval(Grid& grid, DataMatrix& dataset)
      : grid(grid), dataset(dataset), isPrepared(false) {}
  virtual ~OperationMultipleEval() {}

  virtual void mult(DataVector& alpha, DataVector& result) = 0;
  virtual void mult(DataVector& alpha, DataVector& result, size_t startIndexData,
                    size_t endIndexData) {
    throw not_implemented_exception();
  }
  virtual void multTranspose(DataVector& source, DataVector& result) = 0;
  virtual void multTranspose(DataVector& source, DataVector& result, size_t startIndexGrid,
                             size_t endIndexGrid) {
    throw not_implemented_exception();
  }
  void eval(DataVector& alpha, DataVector& result) { this->mult(alpha, result); }
  virtual void prepare() {}
  virtual double getDuration() = 0;
  virtual std::string getImplementationName() {
    throw operation_exception(
        "error: OperationMultipleEval::getImplementationName(): not implemented for this kernel");
  }
};

// y = A x for the solvers (base/src/sgpp/base/operation/hash/OperationMatrix.hpp)
class OperationMatrix {
 public:
  virtual ~OperationMatrix() {}
  virtual void mult(DataVector& alpha, DataVector& result) = 0;
};

}  // namespace base
}  // namespace sgpp
