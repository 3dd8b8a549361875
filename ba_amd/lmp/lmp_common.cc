// Names, logging, fatal errors and platform management of the lmp layer.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <mutex>

#include "la_global.h"
#include "lmp_log.h"
#include "platform_management.h"

namespace hiflow {
namespace la {

std::string platform_name(PLATFORM p) {
  switch (p) {
    case CPU: return "CPU";
    case GPU: return "GPU";
    case OPENCL: return "OPENCL";
    case HIP: return "HIP";
  }
  return "unknown";
}

std::string implementation_name(IMPLEMENTATION i) {
  switch (i) {
    case NAIVE: return "Naive";
    case BLAS: return "BLAS";
    case OPENMP: return "OPENMP";
    case MKL: return "MKL";
    case SCALAR: return "SCALAR";
    case SCALAR_TEX: return "SCALAR_TEX";
    case OPEN_CL: return "OPEN_CL";
    case CUBLAS2: return "CUBLAS2";
    case HYPRE: return "HYPRE";
    case PETSC: return "PETSC";
    case HIPKERNEL: return "HIPKERNEL";
  }
  return "unknown";
}

std::string format_name(MATRIX_FORMAT f) {
  switch (f) {
    case DENSE: return "DENSE";
    case CSR: return "CSR";
    case COO: return "COO";
    case ELL: return "ELL";
  }
  return "unknown";
}

namespace {
FatalHandler g_fatal;
int g_log_level = 0;
std::mutex g_ctx_mutex;
hflmp_ctx *g_ctx = nullptr;
}  // namespace

void set_fatal_handler(FatalHandler h) { g_fatal = std::move(h); }

void fatal_error(const std::string &msg) {
  if (g_fatal) {
    g_fatal(msg);
    throw LmpError(msg);
  }
  std::cerr << "ERROR [lmp] " << msg << std::endl;
  std::exit(-1);
}

void log_info(const std::string &who, const std::string &msg) {
  std::cout << "INFO [" << who << "] " << msg << std::endl;
}

int log_level() { return g_log_level; }
void set_log_level(int level) { g_log_level = level; }

void init_platform(SYSTEM &my_system) {
  if (my_system.Platform != HIP)
    LMP_FATAL("init_platform: platform " << platform_name(my_system.Platform)
                                         << " is not provided by this library (HIP only)");
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  if (g_ctx) {
    int dev = -1;
    hflmp_ctx_device(g_ctx, &dev);
    if (dev == my_system.device) {
      my_system.Initialized = true;
      return;
    }
    hflmp_ctx_destroy(g_ctx);
    g_ctx = nullptr;
  }
  int rc = hflmp_ctx_create(my_system.device, &g_ctx);
  if (rc != HFLMP_OK) {
    g_ctx = nullptr;
    LMP_FATAL("init_platform: no usable HIP device " << my_system.device << " (code " << rc
                                                     << ")");
  }
  my_system.Initialized = true;
}

void stop_platform(SYSTEM &my_system) {
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  if (g_ctx) hflmp_ctx_destroy(g_ctx);
  g_ctx = nullptr;
  my_system.Initialized = false;
}

hflmp_ctx *hip_ctx() {
  {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (g_ctx) return g_ctx;
  }
  SYSTEM sys;
  init_platform(sys);
  return g_ctx;
}

void hip_check(int rc, const char *what) {
  if (rc != HFLMP_OK) {
    const char *msg = g_ctx ? hflmp_ctx_last_error(g_ctx) : "";
    LMP_FATAL(what << " failed (code " << rc << "): " << msg);
  }
}

}  // namespace la
}  // namespace hiflow
