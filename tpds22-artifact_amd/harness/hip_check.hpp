// hip_check.hpp -- exit with the failing HIP call's name and error string.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define HIP_CHECK(call)                                                                  \
  do {                                                                                   \
    const hipError_t hip_check_e_ = (call);                                              \
    if (hip_check_e_ != hipSuccess) {                                                    \
      fprintf(stderr, "%s:%d: %s failed: %s\n", __FILE__, __LINE__, #call,               \
              hipGetErrorString(hip_check_e_));                                          \
      exit(1);                                                                           \
    }                                                                                    \
  } while (0)
