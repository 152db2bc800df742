// Per-thread status recording shared by every translation unit of the library.
#pragma once
#include <hip/hip_runtime.h>

namespace ctlab {
void set_error(int code, const char* fmt, ...);
void clear_error();
// records the error and returns -1 when e != hipSuccess, else returns 0
int hip_check(hipError_t e, const char* what);
}  // namespace ctlab
