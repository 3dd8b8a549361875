// Platform / implementation / format enumerations of the local
// multi-platform linear algebra, with the HIP platform this library adds.
// Interface of linear_algebra/lmp/la_global.h:41-85 (HiFlow3); the values of
// the reference's enumerators are kept so configuration code that stores
// them stays valid, and HIP / HIPKERNEL are appended.
#pragma once

#include <string>

namespace hiflow {
namespace la {

enum PLATFORM { CPU = 0, GPU = 1, OPENCL = 2, HIP = 3 };

enum IMPLEMENTATION {
  NAIVE,
  BLAS,
  OPENMP,
  MKL,
  SCALAR,
  SCALAR_TEX,
  OPEN_CL,
  CUBLAS2,
  HYPRE,
  PETSC,
  HIPKERNEL  // hand-written gfx950 kernels (libhflmp_hip)
};

enum MATRIX_FORMAT { DENSE, CSR, COO, ELL };

// System descriptor (la_global.h:71-85), reduced to what the HIP platform
// uses: the platform and, for HIP, the device ordinal.
struct SYSTEM {
  enum PLATFORM Platform = HIP;
  int device = 0;
  bool Initialized = false;
  int rank = 0;
  bool is_double = true;
};

std::string platform_name(PLATFORM p);
std::string implementation_name(IMPLEMENTATION i);
std::string format_name(MATRIX_FORMAT f);

}  // namespace la
}  // namespace hiflow
