// errors.cpp - thread-local last-error string of the C ABI.
#include "errors.hpp"

#include "elmm_gpu.h"

namespace elmm {
static thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
const char *last_error() { return g_last_error.c_str(); }
}  // namespace elmm

extern "C" const char *elmm_last_error(void) { return elmm::last_error(); }
extern "C" const char *elmm_version(void) { return "elmm-mi355x 0.1 (gfx950)"; }
