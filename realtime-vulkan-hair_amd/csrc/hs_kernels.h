// hs_kernels.h - launch interface between the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device_math.hpp"

namespace hs {

constexpr uint32_t kStepBlock = 256;
constexpr uint32_t kMaxColliders = 8;

// Everything one step launch needs, passed by value as the kernel argument.
struct StepArgs {
    float* pos;           // [3][N][Sp]
    float* vel;           // [3][N][Sp]
    float* corr;          // [3][N][Sp] or null (correction vectors not kept)
    const float* seg;     // [Sp] segment rest length
    const float* noise;   // [Sp] per-strand wind noise or null
    uint32_t S, N, Sp;
    // voxel grid (compute.comp:9-11, 204-206)
    const int4* gprev;    // grid of the previous step (its friction is applied here)
    int4* gnext;          // grid this step splats into (zeroed)
    int4* gclear;         // grid to zero for the next step, or null
    uint32_t G, clear_cells;
    float ox, oy, oz, inv_h;
    float mu_prev, omu_prev;  // friction of the previous step and 1 - friction
    // step constants
    float dt, inv_dt, dt2, D, k, vmax, sphere_r;
    float fx, fy, fz;     // gravity + uniform wind
    float ndx, ndy, ndz;  // per-strand noise direction
    uint32_t ncol;
    Collider col[kMaxColliders];
};

struct UnpackGrid {
    const int4* grid;
    uint32_t G;
    float ox, oy, oz, inv_h, mu, omu;
};

hipError_t launch_step(const StepArgs& a, bool gather, bool splat, hipStream_t st);
hipError_t launch_pack(const float* aos_dev, uint32_t S, uint32_t N, uint32_t Sp, float* pos,
                       float* vel, float* corr, hipStream_t st);
hipError_t launch_unpack(float* aos_dev, uint32_t S, uint32_t N, uint32_t Sp, const float* pos,
                         const float* vel, const float* corr, const UnpackGrid* g, hipStream_t st);

}  // namespace hs
