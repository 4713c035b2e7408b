// rtspatial::Exception (reference include/rtspatial/utils/exception.h:104-117)
// and the check that turns a C-ABI status into it.
#ifndef RTSPATIAL_AMD_UTILS_EXCEPTION_H
#define RTSPATIAL_AMD_UTILS_EXCEPTION_H

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

#include "rts_c.h"

namespace rtspatial {

class Exception : public std::runtime_error {
 public:
  explicit Exception(const std::string& msg) : std::runtime_error(msg) {}
};

inline void check(int status, const char* what) {
  if (status != RTS_OK) throw Exception(std::string(what) + ": " + rts_last_error());
}

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Exception(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace rtspatial
#endif
