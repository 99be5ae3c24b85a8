#pragma once
// Process-wide device context of the C++ layer and error helpers.
#include <string>

#include "../../../include/timeevolver.h"

namespace tecpp {
// The context on device TE_DEVICE (default 0), created on first use. There is no CPU fallback:
// without a usable device the process reports the HIP error and exits.
te_ctx* ctx();
// Reference behaviour for fatal conditions: log FATAL and exit(1).
[[noreturn]] void fatal(const std::string& msg);
// exit(1) with the library's message unless rc == TE_OK.
void check(int rc, const char* what);
}  // namespace tecpp
