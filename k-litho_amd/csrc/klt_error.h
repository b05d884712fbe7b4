// klt_error.h -- the calling thread's last error message (klt_last_error) and
// the status helpers shared by every translation unit of libklt.so; free of
// HIP so that the host-only files build with any C++ compiler.
#pragma once
#include <string>

#include "../../include/klt.h"

namespace klt {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);

}  // namespace klt

#define KLT_TRY(expr)              \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != KLT_OK) return rc_; \
  } while (0)
