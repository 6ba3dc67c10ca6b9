// errors.hpp - thread-local last-error string of the C ABI.
#pragma once
#include <string>

namespace elmm {
// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string &msg);
const char *last_error();
}  // namespace elmm
