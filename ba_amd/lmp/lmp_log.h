// Logging and fatal errors of the lmp layer.  The reference reports a
// failure with LOG_ERROR and ends the program (common/macros.h:79-84,
// lmp/cuda/cuda_utils.h:25-55); that stays the default here.  An embedding
// that must survive errors (the C driver, tests) installs a handler that
// throws instead.
#pragma once

#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>

namespace hiflow {
namespace la {

struct LmpError : std::runtime_error {
  explicit LmpError(const std::string &m) : std::runtime_error(m) {}
};

// Called with the message of a fatal error; the default prints it and calls
// exit(-1).  A handler that returns makes the caller throw LmpError.
using FatalHandler = std::function<void(const std::string &)>;
void set_fatal_handler(FatalHandler h);
[[noreturn]] void fatal_error(const std::string &msg);
void log_info(const std::string &who, const std::string &msg);
int log_level();
void set_log_level(int level);

}  // namespace la
}  // namespace hiflow

#define LMP_FATAL(msg)                                                         \
  do {                                                                         \
    std::ostringstream lmp_os_;                                                \
    lmp_os_ << msg;                                                            \
    ::hiflow::la::fatal_error(lmp_os_.str());                                  \
  } while (0)

#define LMP_INFO(who, msg)                                                     \
  do {                                                                         \
    if (::hiflow::la::log_level() > 0) {                                       \
      std::ostringstream lmp_os_;                                              \
      lmp_os_ << msg;                                                          \
      ::hiflow::la::log_info(who, lmp_os_.str());                              \
    }                                                                          \
  } while (0)
