// Factory of local vectors and matrices (linear_algebra/lmp/init_vec_mat.h:
// 48-82, init_vec_mat.cc:51-482): objects are created only through
// init_vector / init_matrix keyed by platform, implementation and format.
// This library provides the HIP platform (implementation HIPKERNEL, CSR);
// any other key is reported as not provided, the reference's behaviour for
// unsupported combinations (init_vec_mat.cc:191-195).
#pragma once

#include <string>

#include "la_global.h"
#include "lmatrix.h"
#include "lvector.h"

namespace hiflow {
namespace la {

template <typename ValueType>
lVector<ValueType> *init_vector(const int size, const std::string name, const PLATFORM platform,
                                const IMPLEMENTATION implementation);

template <typename ValueType>
lMatrix<ValueType> *init_matrix(const int init_nnz, const int init_num_row,
                                const int init_num_col, const std::string init_name,
                                const PLATFORM platform, const IMPLEMENTATION implementation,
                                const MATRIX_FORMAT matrix_format);

// Configuration strings ("Platform", "Implementation", "MatrixFormat" keys)
// to enumerators; read_LA_parameters in init_vec_mat.cc:51-143.
void read_LA_parameters(const std::string &platform_str, const std::string &impl_str,
                        const std::string &format_str, PLATFORM *platform,
                        IMPLEMENTATION *implementation, MATRIX_FORMAT *format);

}  // namespace la
}  // namespace hiflow
