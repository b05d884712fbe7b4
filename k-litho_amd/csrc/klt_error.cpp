// klt_error.cpp -- per-thread error state of libklt.so.
#include "klt_error.h"

namespace klt {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace klt

extern "C" const char* klt_last_error(void) { return klt::g_last_error.c_str(); }
