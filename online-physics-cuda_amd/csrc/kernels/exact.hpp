// The CPU reference's arithmetic, as device code.
//
// The reference evaluates every expression node over the whole lattice in double and stores each cell-shaped node's
// result as a float (src/field/Law.cpp:612-664, 757-1466), on threads that flush denormals to zero on input and output
// (src/field/Parallel.hpp:18-28: FTZ and DAZ). A kernel that wants the reference's bits computes the same way: doubles
// inside a node, one rounding to float at the node's end, and a float below the smallest normal flushed to a signed zero.
// These helpers are shared by the handwritten kernels and by the code the generator emits (the same text is pasted into
// generated sources, so the two cannot drift apart).
//
// Flushing: every source that includes this is compiled with the f32 denormal mode set to flush (hipcc and hipRTC
// option -fgpu-flush-denormals-to-zero): the hardware then flushes f32 results, conversions to f32 included, and f32
// inputs, to signed zero, as the CPU's FTZ/DAZ threads do, so fe_ftz() is the identity. A build without that mode gets
// the explicit compare-and-select instead (FE_SOFT_FTZ).
#pragma once

#ifdef FE_SOFT_FTZ
#define FE_FTZ_BODY return __builtin_fabsf(f) < 1.17549435e-38f ? __builtin_copysignf(0.0f, f) : f;
#define FE_FTZ_TEXT "{ return __builtin_fabsf(f) < 1.17549435e-38f ? __builtin_copysignf(0.0f, f) : f; }\n"
#else
#define FE_FTZ_BODY return f;
#define FE_FTZ_TEXT "{ return f; }\n"
#endif

#define FE_EXACT_HELPERS                                                                                               \
    __device__ __forceinline__ float fe_ftz(float f) { FE_FTZ_BODY }                                                   \
    __device__ __forceinline__ float fe_rf(double d) { return fe_ftz(static_cast<float>(d)); }                         \
    __device__ __forceinline__ double fe_min(double a, double b) { return b < a ? b : a; }                             \
    __device__ __forceinline__ double fe_max(double a, double b) { return a < b ? b : a; }                             \
    __device__ __forceinline__ double fe_clamp(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

#define FE_EXACT_HELPERS_TEXT                                                                                           \
    "__device__ __forceinline__ float fe_ftz(float f)\n" FE_FTZ_TEXT                                                   \
    "__device__ __forceinline__ float fe_rf(double d) { return fe_ftz(static_cast<float>(d)); }\n"                     \
    "__device__ __forceinline__ double fe_min(double a, double b) { return b < a ? b : a; }\n"                         \
    "__device__ __forceinline__ double fe_max(double a, double b) { return a < b ? b : a; }\n"                         \
    "__device__ __forceinline__ double fe_clamp(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }\n"
