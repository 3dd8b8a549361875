// Platform start/stop (linear_algebra/lmp/platform_management.h:32-35) for
// the HIP platform: init_platform creates the process's device context
// (device, stream, reduction scratch) through the C ABI, stop_platform
// releases it.  hip_ctx() hands the context to the HIP classes.
#pragma once

#include "hflmp_hip.h"
#include "la_global.h"

namespace hiflow {
namespace la {

void init_platform(SYSTEM &my_system);
void stop_platform(SYSTEM &my_system);

// The context of the HIP platform; created on device 0 at first use when
// init_platform was not called.  A failure to create it is fatal.
hflmp_ctx *hip_ctx();
// Fatal error with the context's message when rc != HFLMP_OK.
void hip_check(int rc, const char *what);

}  // namespace la
}  // namespace hiflow
