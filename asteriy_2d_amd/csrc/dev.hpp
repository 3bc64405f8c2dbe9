// dev.hpp -- the device-function qualifier shared by the kernels' headers.
#pragma once
#include <hip/hip_runtime.h>

#define AST_DEV __device__ __forceinline__
