// Linear-algebra descriptors handed to the solver templates
// (linear_algebra/la_descriptor.h:58-72).
#pragma once

#include "coupled_matrix.h"
#include "coupled_vector.h"

namespace hiflow {
namespace la {

class LADescriptorCoupledD {
 public:
  typedef CoupledMatrix<double> MatrixType;
  typedef CoupledVector<double> VectorType;
  typedef double DataType;
};

class LADescriptorCoupledS {
 public:
  typedef CoupledMatrix<float> MatrixType;
  typedef CoupledVector<float> VectorType;
  typedef float DataType;
};

}  // namespace la
}  // namespace hiflow
