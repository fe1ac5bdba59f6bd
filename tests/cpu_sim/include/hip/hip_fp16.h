// CPU simulation of the half-precision conversions the kernels use.
#pragma once
#include <hip/hip_runtime.h>

typedef _Float16 __half;
inline float __half2float(__half h) { return (float)h; }
inline __half __float2half(float f) { return (__half)f; }
inline unsigned short __half_as_ushort(__half h)
{
    unsigned short u;
    std::memcpy(&u, &h, 2);
    return u;
}
inline __half __ushort_as_half(unsigned short u)
{
    __half h;
    std::memcpy(&h, &u, 2);
    return h;
}
