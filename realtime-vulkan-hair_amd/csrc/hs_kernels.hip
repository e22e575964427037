// hs_kernels.hip - HIP kernels of the strand-physics pass for gfx950.
//
// State layout in HBM (strand-major SoA, one plane per component):
//   pos[c][i][s], vel[c][i][s], c in {x,y,z}, i = point 0..N-1, s = strand
//   0..Sp-1 (Sp = S rounded up to 64).  At a fixed point index the 64 lanes of
//   a wavefront read 64 consecutive floats of each plane: one 256 B line.
// One lane owns one strand and sweeps its points root to tip, keeping only
// the parent position and the previous point's pending velocity in registers
// (the DFTL correction of point i-1 needs point i's correction vector).
//
// The reference's step (compute.comp:124-303) has a global dependency inside
// it: every strand's splat must land before any strand gathers.  The kernels
// here are step-skewed: launch k applies step k-1's grid friction to each
// point as it is loaded (gather from grid[k-1]) and then runs step k's
// predict / FTL / DFTL / splat (into grid[k]).  Positions written by launch k
// are the positions after step k; stored velocities are the post-DFTL
// velocities of step k, whose friction is applied by launch k+1 or by the
// read-back kernel.  Per point the operations and their order are exactly
// those of the unskewed step, so results are bit-identical.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device_math.hpp"
#include "hs_kernels.h"

namespace hs {

// ---------------------------------------------------------------------------
// Grid helpers
// ---------------------------------------------------------------------------

// Grid velocity at grid coordinate q (compute.comp:264-293): sum over the
// touched cells with density > 0 of (w * (1/float(density))) * vec3(vel),
// cells visited x-major, then y, then z as the shader's a/b/c loops.
__device__ __forceinline__ f3 grid_gather(const int4* __restrict__ g, f3 q, int G)
{
    int x0, x1, y0, y1, z0, z1;
    cell_range(q.x, G, x0, x1);
    cell_range(q.y, G, y0, y1);
    cell_range(q.z, G, z0, z1);
    f3 gv = mk(0.0f, 0.0f, 0.0f);
    for (int a = x0; a <= x1; ++a) {
        const float wx = clamp01(1.0f - fabsf(q.x - (float)a));
        for (int b = y0; b <= y1; ++b) {
            const float wy = clamp01(1.0f - fabsf(q.y - (float)b));
            const float wxy = wx * wy;
            for (int c = z0; c <= z1; ++c) {
                const int4 cell = g[(size_t)a + (size_t)b * G + (size_t)c * G * G];
                if (cell.w > 0) {
                    const float wz = clamp01(1.0f - fabsf(q.z - (float)c));
                    const float w = wxy * wz;
                    const float t = w * (1.0f / (float)cell.w);
                    gv = add(gv, smul(t, mk((float)cell.x, (float)cell.y, (float)cell.z)));
                }
            }
        }
    }
    return gv;
}

// Trilinear fixed-point splat at grid coordinate q (compute.comp:224-252).
__device__ __forceinline__ void grid_splat(int4* __restrict__ g, f3 q, f3 v, int G)
{
    const float SCALE = 1000000.0f;
    int x0, x1, y0, y1, z0, z1;
    cell_range(q.x, G, x0, x1);
    cell_range(q.y, G, y0, y1);
    cell_range(q.z, G, z0, z1);
    for (int a = x0; a <= x1; ++a) {
        const float wx = clamp01(1.0f - fabsf(q.x - (float)a));
        for (int b = y0; b <= y1; ++b) {
            const float wy = clamp01(1.0f - fabsf(q.y - (float)b));
            const float wxy = wx * wy;
            for (int c = z0; c <= z1; ++c) {
                const float wz = clamp01(1.0f - fabsf(q.z - (float)c));
                const float w = wxy * wz;
                const f3 wv = smul(w, v);
                int* cell = reinterpret_cast<int*>(g + ((size_t)a + (size_t)b * G + (size_t)c * G * G));
                __hip_atomic_fetch_add(cell + 0, sat_i32(SCALE * wv.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(cell + 1, sat_i32(SCALE * wv.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(cell + 2, sat_i32(SCALE * wv.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(cell + 3, sat_i32(SCALE * w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

__device__ __forceinline__ f3 grid_coord(f3 p, f3 origin, float inv_h)
{
    return muls(sub(p, origin), inv_h);
}

// Friction blend (compute.comp:296-297): (1 - mu) * v + mu * v_grid.
__device__ __forceinline__ f3 friction(f3 v, f3 gv, float mu, float omu)
{
    return add(smul(omu, v), smul(mu, gv));
}

// ---------------------------------------------------------------------------
// Collider penalty (compute.comp:155-184)
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 collide(f3 x, f3 force, const StepArgs& a)
{
    f3 acc = mk(0.0f, 0.0f, 0.0f);
    int hits = 0;
    for (uint32_t j = 0; j < a.ncol; ++j) {
        const float* T = a.col[j].m;
        if (j == 0) {
            // sphere: compute.comp:160-169 (centre = transform[3].xyz)
            const f3 dv = sub(x, mk(T[12], T[13], T[14]));
            const float dist = length(dv);
            if (dist < a.sphere_r) {
                const float d = a.sphere_r - dist;
                const f3 n = muls(dv, 1.0f / dist);
                acc = add(acc, smul(a.k * d, n));
                hits += 1;
            }
        } else {
            // ellipsoid: compute.comp:64-79, 170-179
            const f3 tp = mat_vec(T + 16, x, 1.0f);
            if (length(tp) <= 1.0f) {
                const f3 surf = mat_vec(T, normalize(tp), 1.0f);
                const float d = length(sub(surf, x));
                const f3 n = normalize(normalize(mat_vec(T + 32, tp, 0.0f)));
                acc = add(acc, smul(a.k * d, n));
                hits += 1;
            }
        }
    }
    if (hits > 0) {
        const float m = (float)hits;
        force = add(force, mk(acc.x / m, acc.y / m, acc.z / m));
    }
    return force;
}

// ---------------------------------------------------------------------------
// The step kernel
// ---------------------------------------------------------------------------
template <bool GATHER, bool SPLAT>
__global__ __launch_bounds__(kStepBlock) void k_step(StepArgs a)
{
    const uint32_t s = blockIdx.x * kStepBlock + threadIdx.x;
    const uint32_t Sp = a.Sp;
    const uint32_t N = a.N;
    const size_t plane = (size_t)N * Sp;
    float* __restrict__ px = a.pos;
    float* __restrict__ py = a.pos + plane;
    float* __restrict__ pz = a.pos + 2 * plane;
    float* __restrict__ vx = a.vel;
    float* __restrict__ vy = a.vel + plane;
    float* __restrict__ vz = a.vel + 2 * plane;
    const f3 origin = mk(a.ox, a.oy, a.oz);
    const int G = (int)a.G;

    if (s < a.S) {
        f3 par = mk(px[s], py[s], pz[s]);
        const float r = a.seg[s];
        f3 fs = mk(a.fx, a.fy, a.fz);
        if (a.noise) fs = add(fs, smul(a.noise[s], mk(a.ndx, a.ndy, a.ndz)));

        f3 vpend = mk(0.0f, 0.0f, 0.0f);   // velocity of point i-1 awaiting DFTL
        f3 ppend = mk(0.0f, 0.0f, 0.0f);   // position of point i-1
        for (uint32_t i = 1; i < N; ++i) {
            const size_t o = (size_t)i * Sp + s;
            const f3 x = mk(px[o], py[o], pz[o]);
            f3 u = mk(vx[o], vy[o], vz[o]);
            if (GATHER) {
                // previous step's phase 3 (compute.comp:257-298) for this point
                const f3 gv = grid_gather(a.gprev, grid_coord(x, origin, a.inv_h), G);
                u = friction(u, gv, a.mu_prev, a.omu_prev);
            }
            const f3 f = collide(x, fs, a);
            // predict (compute.comp:187)
            const f3 pred = add(add(x, smul(a.dt, u)), smul(a.dt2, f));
            // follow the leader (compute.comp:191-192)
            const f3 pn = add(par, smul(r, normalize(sub(pred, par))));
            // damped velocity, clamp, correction (compute.comp:195-201)
            f3 vel = smul(a.D, muls(sub(pn, x), a.inv_dt));
            if (length(vel) > a.vmax) vel = muls(normalize(vel), a.vmax);
            const f3 corr = smul(a.D, sub(pn, pred));
            px[o] = pn.x;
            py[o] = pn.y;
            pz[o] = pn.z;
            if (a.corr) {
                a.corr[o] = corr.x;
                a.corr[plane + o] = corr.y;
                a.corr[2 * plane + o] = corr.z;
            }
            if (i >= 2) {
                // DFTL for point i-1 (compute.comp:213-215), then its splat
                vpend = sub(vpend, muls(corr, a.inv_dt));
                const size_t op = o - Sp;
                vx[op] = vpend.x;
                vy[op] = vpend.y;
                vz[op] = vpend.z;
                if (SPLAT) grid_splat(a.gnext, grid_coord(ppend, origin, a.inv_h), vpend, G);
            }
            vpend = vel;
            ppend = pn;
            par = pn;
        }
        // tip: no correction
        const size_t ot = (size_t)(N - 1) * Sp + s;
        vx[ot] = vpend.x;
        vy[ot] = vpend.y;
        vz[ot] = vpend.z;
        if (SPLAT) grid_splat(a.gnext, grid_coord(ppend, origin, a.inv_h), vpend, G);
    }

    // Clear the grid buffer the next launch will splat into.
    if (a.gclear) {
        const uint32_t nthreads = gridDim.x * kStepBlock;
        const uint32_t t = blockIdx.x * kStepBlock + threadIdx.x;
        for (uint32_t c = t; c < a.clear_cells; c += nthreads) a.gclear[c] = make_int4(0, 0, 0, 0);
    }
}

// ---------------------------------------------------------------------------
// AoS <-> SoA (reference Strand layout at the boundary)
// ---------------------------------------------------------------------------
// One thread per (strand, point).
__global__ void k_pack(const float4* __restrict__ aos, uint32_t S, uint32_t N, uint32_t Sp,
                       float* __restrict__ pos, float* __restrict__ vel, float* __restrict__ corr)
{
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)S * N) return;
    const uint32_t s = (uint32_t)(tid / N);
    const uint32_t i = (uint32_t)(tid % N);
    const size_t plane = (size_t)N * Sp;
    const size_t base = (size_t)s * 3 * N;
    const float4 p = aos[base + i];
    const float4 v = aos[base + N + i];
    const size_t o = (size_t)i * Sp + s;
    pos[o] = p.x;
    pos[plane + o] = p.y;
    pos[2 * plane + o] = p.z;
    vel[o] = v.x;
    vel[plane + o] = v.y;
    vel[2 * plane + o] = v.z;
    if (corr) {
        const float4 c = aos[base + 2 * N + i];
        corr[o] = c.x;
        corr[plane + o] = c.y;
        corr[2 * plane + o] = c.z;
    }
}

// SoA -> AoS; when `g` is given, applies the pending grid friction to
// velocities of points 1..N-1 (compute.comp:257-298) on the way out.
__global__ void k_unpack(float4* __restrict__ aos, uint32_t S, uint32_t N, uint32_t Sp,
                         const float* __restrict__ pos, const float* __restrict__ vel,
                         const float* __restrict__ corr, const int4* __restrict__ g, uint32_t G,
                         float ox, float oy, float oz, float inv_h, float mu, float omu)
{
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)S * N) return;
    const uint32_t s = (uint32_t)(tid / N);
    const uint32_t i = (uint32_t)(tid % N);
    const size_t plane = (size_t)N * Sp;
    const size_t base = (size_t)s * 3 * N;
    const size_t o = (size_t)i * Sp + s;
    const f3 p = mk(pos[o], pos[plane + o], pos[2 * plane + o]);
    f3 v = mk(vel[o], vel[plane + o], vel[2 * plane + o]);
    if (g && i >= 1) {
        const f3 gv = grid_gather(g, grid_coord(p, mk(ox, oy, oz), inv_h), (int)G);
        v = friction(v, gv, mu, omu);
    }
    aos[base + i] = make_float4(p.x, p.y, p.z, 1.0f);
    aos[base + N + i] = make_float4(v.x, v.y, v.z, 0.0f);
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (corr) c = make_float4(corr[o], corr[plane + o], corr[2 * plane + o], 0.0f);
    aos[base + 2 * N + i] = c;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
hipError_t launch_step(const StepArgs& a, bool gather, bool splat, hipStream_t st)
{
    const uint32_t blocks = (a.Sp + kStepBlock - 1) / kStepBlock;
    if (gather) {
        if (splat)
            hipLaunchKernelGGL((k_step<true, true>), dim3(blocks), dim3(kStepBlock), 0, st, a);
        else
            hipLaunchKernelGGL((k_step<true, false>), dim3(blocks), dim3(kStepBlock), 0, st, a);
    } else {
        if (splat)
            hipLaunchKernelGGL((k_step<false, true>), dim3(blocks), dim3(kStepBlock), 0, st, a);
        else
            hipLaunchKernelGGL((k_step<false, false>), dim3(blocks), dim3(kStepBlock), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_pack(const float* aos_dev, uint32_t S, uint32_t N, uint32_t Sp, float* pos,
                       float* vel, float* corr, hipStream_t st)
{
    const size_t n = (size_t)S * N;
    const uint32_t bs = 256;
    const size_t blocks = (n + bs - 1) / bs;
    hipLaunchKernelGGL(k_pack, dim3((uint32_t)blocks), dim3(bs), 0, st,
                       reinterpret_cast<const float4*>(aos_dev), S, N, Sp, pos, vel, corr);
    return hipGetLastError();
}

hipError_t launch_unpack(float* aos_dev, uint32_t S, uint32_t N, uint32_t Sp, const float* pos,
                         const float* vel, const float* corr, const UnpackGrid* g, hipStream_t st)
{
    const size_t n = (size_t)S * N;
    const uint32_t bs = 256;
    const size_t blocks = (n + bs - 1) / bs;
    if (g)
        hipLaunchKernelGGL(k_unpack, dim3((uint32_t)blocks), dim3(bs), 0, st,
                           reinterpret_cast<float4*>(aos_dev), S, N, Sp, pos, vel, corr, g->grid, g->G,
                           g->ox, g->oy, g->oz, g->inv_h, g->mu, g->omu);
    else
        hipLaunchKernelGGL(k_unpack, dim3((uint32_t)blocks), dim3(bs), 0, st,
                           reinterpret_cast<float4*>(aos_dev), S, N, Sp, pos, vel, corr,
                           (const int4*)nullptr, 0u, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f);
    return hipGetLastError();
}

}  // namespace hs
