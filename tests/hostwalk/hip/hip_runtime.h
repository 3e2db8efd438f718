// Test harness only: lets tests/hostwalk/walk_host.cpp compile the device
// ray-walk header (yrt-pet_amd/csrc/yrtmi_rays.hpp) with the host compiler,
// so the walk's logic is checked on the CPU.  It defines the HIP function
// qualifiers away and gives host versions of the few device intrinsics the
// header uses (correctly rounded FP32 ops, exp, bit casts, atomic add).  Nothing in the
// library includes it.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(x)
struct float3
{
	float x, y, z;
};
inline float3 make_float3(float a, float b, float c) { return {a, b, c}; }
struct float4
{
	float x, y, z, w;
};
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline float __int_as_float(int v)
{
	float f;
	std::memcpy(&f, &v, 4);
	return f;
}
inline float __uint_as_float(uint32_t v)
{
	float f;
	std::memcpy(&f, &v, 4);
	return f;
}
inline int __float_as_int(float f)
{
	int v;
	std::memcpy(&v, &f, 4);
	return v;
}
inline uint32_t __float_as_uint(float f)
{
	uint32_t v;
	std::memcpy(&v, &f, 4);
	return v;
}
inline float __fadd_rn(float a, float b) { return a + b; }
inline float __fsub_rn(float a, float b) { return a - b; }
inline float __fmul_rn(float a, float b) { return a * b; }
inline float __fdiv_rn(float a, float b) { return a / b; }
inline float __fsqrt_rn(float a) { return std::sqrt(a); }
inline float __expf(float a) { return std::exp(a); }
#define __builtin_amdgcn_exp2f(x) std::exp2(x)
inline void unsafeAtomicAdd(float* p, float v) { *p += v; }
using std::max;
using std::min;
typedef int hipError_t;
typedef void* hipStream_t;
