// CPU simulation: hipExtLaunchKernelGGL is defined in hip_runtime.h.
#pragma once
#include <hip/hip_runtime.h>
