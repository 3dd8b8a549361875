#include "init_vec_mat.h"

#include "lmatrix_csr_hip.h"
#include "lmp_log.h"
#include "lvector_hip.h"

namespace hiflow {
namespace la {

template <typename T>
lVector<T> *init_vector(const int size, const std::string name, const PLATFORM platform,
                        const IMPLEMENTATION implementation) {
  if (platform == HIP && (implementation == HIPKERNEL || implementation == NAIVE))
    return new HIP_lVector<T>(size, name);
  LMP_FATAL("init_vector: platform " << platform_name(platform) << " / implementation "
                                     << implementation_name(implementation)
                                     << " is not provided by this library");
}

template <typename T>
lMatrix<T> *init_matrix(const int init_nnz, const int init_num_row, const int init_num_col,
                        const std::string init_name, const PLATFORM platform,
                        const IMPLEMENTATION implementation, const MATRIX_FORMAT matrix_format) {
  if (platform == HIP && (implementation == HIPKERNEL || implementation == NAIVE) &&
      matrix_format == CSR)
    return new HIP_CSR_lMatrix<T>(init_nnz, init_num_row, init_num_col, init_name);
  LMP_FATAL("init_matrix: platform " << platform_name(platform) << " / implementation "
                                     << implementation_name(implementation) << " / format "
                                     << format_name(matrix_format)
                                     << " is not provided by this library");
}

void read_LA_parameters(const std::string &platform_str, const std::string &impl_str,
                        const std::string &format_str, PLATFORM *platform,
                        IMPLEMENTATION *implementation, MATRIX_FORMAT *format) {
  if (platform_str == "HIP") *platform = HIP;
  else if (platform_str == "CPU") *platform = CPU;
  else if (platform_str == "GPU") *platform = GPU;
  else if (platform_str == "OPENCL") *platform = OPENCL;
  else LMP_FATAL("read_LA_parameters: unknown platform " << platform_str);

  if (impl_str == "HIPKERNEL" || impl_str == "HIP") *implementation = HIPKERNEL;
  else if (impl_str == "Naive") *implementation = NAIVE;
  else if (impl_str == "BLAS") *implementation = BLAS;
  else if (impl_str == "OPENMP") *implementation = OPENMP;
  else if (impl_str == "MKL") *implementation = MKL;
  else if (impl_str == "SCALAR") *implementation = SCALAR;
  else if (impl_str == "SCALAR_TEX") *implementation = SCALAR_TEX;
  else if (impl_str == "OPEN_CL") *implementation = OPEN_CL;
  else if (impl_str == "CUBLAS2") *implementation = CUBLAS2;
  else LMP_FATAL("read_LA_parameters: unknown implementation " << impl_str);

  if (format_str == "CSR") *format = CSR;
  else if (format_str == "COO") *format = COO;
  else if (format_str == "DENSE") *format = DENSE;
  else if (format_str == "ELL") *format = ELL;
  else LMP_FATAL("read_LA_parameters: unknown matrix format " << format_str);
}

template lVector<double> *init_vector<double>(const int, const std::string, const PLATFORM,
                                              const IMPLEMENTATION);
template lVector<float> *init_vector<float>(const int, const std::string, const PLATFORM,
                                            const IMPLEMENTATION);
template lMatrix<double> *init_matrix<double>(const int, const int, const int, const std::string,
                                              const PLATFORM, const IMPLEMENTATION,
                                              const MATRIX_FORMAT);
template lMatrix<float> *init_matrix<float>(const int, const int, const int, const std::string,
                                            const PLATFORM, const IMPLEMENTATION,
                                            const MATRIX_FORMAT);

}  // namespace la
}  // namespace hiflow
