// Host-only stand-in for <hip/hip_runtime.h>: lets the device solver headers
// (hydrocode-fusion_amd/csrc/*.hpp) compile as plain C++ for CPU checks.
#pragma once
#include <math.h>
#include <stdint.h>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __constant__
