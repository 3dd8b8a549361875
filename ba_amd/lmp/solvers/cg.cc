// Instantiations of the lmp-local cg() for double and float
// (linear_algebra/lmp/solvers/cg.cc:35-53).
#include "cg.h"

namespace hiflow {
namespace la {

template <typename ValueType>
int cg(lVector<ValueType> *x, const lVector<ValueType> *b, const lMatrix<ValueType> *matrix,
       const ValueType rel_eps, const ValueType abs_eps, int &max_iter, const int print,
       lPreconditioner<ValueType> *lPrecond) {
  return CGFunctor<ValueType, false>()(x, b, matrix, rel_eps, abs_eps, max_iter, print,
                                       lPrecond);
}

template int cg(lVector<double> *x, const lVector<double> *b, const lMatrix<double> *matrix,
                const double rel_eps, const double abs_eps, int &max_iter, const int print,
                lPreconditioner<double> *lPrecond);
template int cg(lVector<float> *x, const lVector<float> *b, const lMatrix<float> *matrix,
                const float rel_eps, const float abs_eps, int &max_iter, const int print,
                lPreconditioner<float> *lPrecond);

}  // namespace la
}  // namespace hiflow
