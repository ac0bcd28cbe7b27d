// Exception types of the operator boundary, named as in SG++
// (base/src/sgpp/base/exception/*.hpp): the C-ABI maps each to a status
// code and the C++ wrapper re-throws the matching type.
#pragma once

#include <stdexcept>
#include <string>

namespace sgpp {
namespace base {

struct sgpp_exception : public std::runtime_error {
  explicit sgpp_exception(const std::string& m) : std::runtime_error(m) {}
};
struct factory_exception : public sgpp_exception {
  explicit factory_exception(const std::string& m = "factory_exception")
      : sgpp_exception(m) {}
};
struct operation_exception : public sgpp_exception {
  explicit operation_exception(const std::string& m = "operation_exception")
      : sgpp_exception(m) {}
};
struct not_implemented_exception : public sgpp_exception {
  explicit not_implemented_exception(const std::string& m = "not implemented")
      : sgpp_exception(m) {}
};
struct data_exception : public sgpp_exception {
  explicit data_exception(const std::string& m = "data_exception") : sgpp_exception(m) {}
};
struct generation_exception : public sgpp_exception {
  explicit generation_exception(const std::string& m = "generation_exception")
      : sgpp_exception(m) {}
};
struct solver_exception : public sgpp_exception {
  explicit solver_exception(const std::string& m = "solver_exception") : sgpp_exception(m) {}
};
struct application_exception : public sgpp_exception {
  explicit application_exception(const std::string& m = "application_exception")
      : sgpp_exception(m) {}
};

}  // namespace base
}  // namespace sgpp
