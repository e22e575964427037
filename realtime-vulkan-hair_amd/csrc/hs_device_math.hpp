// hs_device_math.hpp - per-point arithmetic of the strand step, device side.
//
// The operation order is fixed so that results are bit-identical to a CPU
// binary32 evaluation of the same expressions (the library is compiled with
// -ffp-contract=off and hipcc's default correctly rounded fp32 division and
// sqrt):
//   dot(a,b) = (ax*bx + ay*by) + az*bz, length = sqrt(dot), normalize(v) =
//   v * (1/length(v)), division by dt or by the voxel size h is a product
//   with the host-rounded reciprocal, mat4*vec4 rows are
//   ((m0*x + m1*y) + m2*z) + m3*w, clamp uses IEEE min/max, float->int is a
//   saturating truncation with NaN -> 0, grid sums wrap modulo 2^32.
// Each helper names the lines of src/shaders/compute.comp it implements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hs {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// scalar * vector (scalar on the left, as the GLSL source writes it)
__device__ __forceinline__ f3 smul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
// vector * scalar
__device__ __forceinline__ f3 muls(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) { return muls(a, 1.0f / length(a)); }
__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// mat4 (column-major m[c*4+r]) * vec4(v, w), xyz part.
__device__ __forceinline__ f3 mat_vec(const float* m, f3 v, float w)
{
    float o[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
        o[r] = ((m[r] * v.x + m[4 + r] * v.y) + m[8 + r] * v.z) + m[12 + r] * w;
    return f3{o[0], o[1], o[2]};
}

// int(f) of the contract: truncation toward zero, saturating to
// [INT32_MIN, INT32_MAX], NaN -> 0.  That is exactly what V_CVT_I32_F32 does
// (out-of-range inputs clamp, NaN converts to 0), so the instruction is used
// directly; a C cast would be undefined for the out-of-range cases.
__device__ __forceinline__ int32_t sat_i32(float f)
{
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}

// Cells one grid coordinate touches (compute.comp:224-229):
// [max(int(floor q), 0), min(int(floor q) + 1, G - 1)].  Computed on the
// floored float so no int overflow can occur for any input.
__device__ __forceinline__ void cell_range(float q, int G, int& lo, int& hi)
{
    float f = floorf(q);
    if (!(f == f)) f = 0.0f;  // NaN -> int() == 0
    // Clamped to [-2, G] without branches: below -2 the range is (0, -1) and
    // above G it is (G, G - 1), both empty, exactly as for f = -2 and f = G.
    f = fminf(fmaxf(f, -2.0f), (float)G);
    const int fi = (int)f;
    lo = fi > 0 ? fi : 0;
    hi = (fi + 1) < (G - 1) ? (fi + 1) : (G - 1);
}

// Per-dispatch constants of one groom's step.
struct Collider {
    float m[48];   // transform, inv, invTrans (Scene.h:23-39)
};

}  // namespace hs
