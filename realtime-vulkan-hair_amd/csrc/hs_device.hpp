// hs_device.hpp - device helpers shared by the step kernels (one build per
// tile shape, hs_step.hip) and the other kernels (hs_kernels.hip): voxel
// ranges, LDS gathers and splats, the start-of-step force, and the access to
// pointers read from the batched launches' argument blob.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device_math.hpp"
#include "hs_forcing.hpp"
#include "hs_kernels.h"

namespace hs {

// floor(x / d) for x in [0, 2^16) and a workgroup-uniform d in [1, 2^16),
// without a division per x.  For d >= 2, m = floor((2^32 - 1) / d) + 1 =
// 2^32/d + e with 0 <= e <= 1 (e = 0 when d is a power of two) fits 32 bits,
// and x*m / 2^32 exceeds x/d by less than x / 2^32 < 2^-16 < 1/d, which never
// carries x/d past the next integer: the high word of x*m is floor(x/d).
// d = 1 (m would be 2^32) is selected apart.
__device__ __forceinline__ uint32_t div_magic(uint32_t d) { return d > 1 ? 0xffffffffu / d + 1u : 0u; }
__device__ __forceinline__ uint32_t div_small(uint32_t x, uint32_t d, uint32_t m)
{
    return d > 1 ? __umulhi(x, m) : x;
}

// ---------------------------------------------------------------------------
// Voxel ranges
// ---------------------------------------------------------------------------
struct Rng {
    int x0, x1, y0, y1, z0, z1;  // inclusive; empty when any lo > hi
};

__device__ __forceinline__ f3 grid_coord(f3 p, f3 origin, float inv_h) { return muls(sub(p, origin), inv_h); }

__device__ __forceinline__ Rng point_range(f3 q, int G)
{
    Rng r;
    cell_range(q.x, G, r.x0, r.x1);
    cell_range(q.y, G, r.y0, r.y1);
    cell_range(q.z, G, r.z0, r.z1);
    return r;
}

// The range tests combine their compares with non-short-circuit '|' and '&':
// plain mask arithmetic instead of a branch per compare.
__device__ __forceinline__ bool rng_empty(const Rng& r) { return (r.x0 > r.x1) | (r.y0 > r.y1) | (r.z0 > r.z1); }

__device__ __forceinline__ bool rng_full(const Rng& r)
{
    return (r.x1 == r.x0 + 1) & (r.y1 == r.y0 + 1) & (r.z1 == r.z0 + 1);
}

// A full 2x2x2 block packed in 31 bits (base voxel, 10 bits per axis, G <= 1024);
// 0 marks "not a full block".
__device__ __forceinline__ uint32_t pack_block(const Rng& r)
{
    return rng_full(r) ? (0x80000000u | (uint32_t)r.x0 | ((uint32_t)r.y0 << 10) | ((uint32_t)r.z0 << 20)) : 0u;
}

__device__ __forceinline__ Rng unpack_block(uint32_t p)
{
    Rng r;
    r.x0 = (int)(p & 1023u);
    r.y0 = (int)((p >> 10) & 1023u);
    r.z0 = (int)((p >> 20) & 1023u);
    r.x1 = r.x0 + 1;
    r.y1 = r.y0 + 1;
    r.z1 = r.z0 + 1;
    return r;
}

// Friction blend (compute.comp:296-297): (1 - mu) * v + mu * v_grid.
__device__ __forceinline__ f3 friction(f3 v, f3 gv, float mu, float omu) { return add(smul(omu, v), smul(mu, gv)); }

// A window of voxels staged in LDS: cell (a,b,c) at ((c-z0)*ey + (b-y0))*ex + (a-x0).
struct Win {
    int x0, y0, z0, ex, ey, ez;
};

__device__ __forceinline__ bool in_win(const Rng& r, const Win& w)
{
    return (r.x0 >= w.x0) & (r.x1 < w.x0 + w.ex) & (r.y0 >= w.y0) & (r.y1 < w.y0 + w.ey) & (r.z0 >= w.z0) &
           (r.z1 < w.z0 + w.ez);
}

__device__ __forceinline__ int win_index(const Win& w, int a, int b, int c)
{
    return ((c - w.z0) * w.ey + (b - w.y0)) * w.ex + (a - w.x0);
}


// 32-bit byte-offset accessors of the [3][N][Sp] planes: x at off, y at
// off + P4, z at off + P8.
__device__ __forceinline__ f3 ld3(const float* base, uint32_t off, uint32_t P4, uint32_t P8)
{
    const char* b = reinterpret_cast<const char*>(base);
    return mk(*reinterpret_cast<const float*>(b + off), *reinterpret_cast<const float*>(b + (off + P4)),
              *reinterpret_cast<const float*>(b + (off + P8)));
}

__device__ __forceinline__ void st3(float* base, uint32_t off, uint32_t P4, uint32_t P8, f3 v)
{
    char* b = reinterpret_cast<char*>(base);
    *reinterpret_cast<float*>(b + off) = v.x;
    *reinterpret_cast<float*>(b + (off + P4)) = v.y;
    *reinterpret_cast<float*>(b + (off + P8)) = v.z;
}

// ---------------------------------------------------------------------------
// Grid gathers.  Cells are visited x-major, then y, then z as the shader's
// a/b/c loops; a float cell holds (vx, vy, vz, 1/density) with 1/density = 0
// for density <= 0, so a skipped cell adds +-0, which never changes the sum
// (the sum starts at +0 and round-to-nearest never produces -0 from x + (-x)).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gather_cell(f3& gv, float4 c, float w)
{
    const float t = w * c.w;
    gv = add(gv, smul(t, mk(c.x, c.y, c.z)));
}

__device__ __forceinline__ float wgt(float q, int a) { return clamp01(1.0f - fabsf(q - (float)a)); }

// Full 2x2x2 block inside an LDS window of float cells.
__device__ __forceinline__ f3 gather_block(const float4* win, const Win& wd, const Rng& r, f3 q)
{
    const float wx0 = wgt(q.x, r.x0), wx1 = wgt(q.x, r.x1);
    const float wy0 = wgt(q.y, r.y0), wy1 = wgt(q.y, r.y1);
    const float wz0 = wgt(q.z, r.z0), wz1 = wgt(q.z, r.z1);
    const int sy = wd.ex, sz = wd.ex * wd.ey;
    const float4* b = win + win_index(wd, r.x0, r.y0, r.z0);
    const float4 c000 = b[0], c001 = b[sz], c010 = b[sy], c011 = b[sy + sz];
    const float4 c100 = b[1], c101 = b[1 + sz], c110 = b[1 + sy], c111 = b[1 + sy + sz];
    f3 gv = mk(0.0f, 0.0f, 0.0f);
    float wxy = wx0 * wy0;
    gather_cell(gv, c000, wxy * wz0);
    gather_cell(gv, c001, wxy * wz1);
    wxy = wx0 * wy1;
    gather_cell(gv, c010, wxy * wz0);
    gather_cell(gv, c011, wxy * wz1);
    wxy = wx1 * wy0;
    gather_cell(gv, c100, wxy * wz0);
    gather_cell(gv, c101, wxy * wz1);
    wxy = wx1 * wy1;
    gather_cell(gv, c110, wxy * wz0);
    gather_cell(gv, c111, wxy * wz1);
    return gv;
}

// Any non-empty block (clipped at the grid boundary, or outside the window):
// from the LDS window when `lds`, else from the global float grid.  The two
// sources stay in separate loops: a pointer selected between LDS and global
// memory would be a generic pointer.
template <bool LDS>
__device__ __forceinline__ f3 gather_any_src(const float4* win, const Win& wd, const float4* __restrict__ g,
                                             const Rng& r, f3 q, int G)
{
    f3 gv = mk(0.0f, 0.0f, 0.0f);
#pragma unroll 1
    for (int a = r.x0; a <= r.x1; ++a) {
        const float wxa = wgt(q.x, a);
#pragma unroll 1
        for (int b = r.y0; b <= r.y1; ++b) {
            const float wxy = wxa * wgt(q.y, b);
#pragma unroll 1
            for (int c = r.z0; c <= r.z1; ++c) {
                const float w = wxy * wgt(q.z, c);
                float4 cell;
                if (LDS)
                    cell = win[win_index(wd, a, b, c)];
                else
                    cell = g[(uint32_t)a + (uint32_t)b * (uint32_t)G + (uint32_t)c * (uint32_t)G * (uint32_t)G];
                gather_cell(gv, cell, w);
            }
        }
    }
    return gv;
}

__device__ __forceinline__ f3 gather_any(const float4* win, const Win& wd, bool lds, const float4* __restrict__ g,
                                         const Rng& r, f3 q, int G)
{
    return lds ? gather_any_src<true>(win, wd, g, r, q, G) : gather_any_src<false>(win, wd, g, r, q, G);
}

// ---------------------------------------------------------------------------
// Splat.  Contributions int(SCALE*(w*v)) and int(SCALE*w) (compute.comp:241-248)
// accumulate in an LDS window holding `nc` copies of every voxel (nc a power
// of two, as many as the window's cells leave room for, at most kMaxCopies);
// lane l adds into copy l & (nc-1).  Hair is dense: tens of lanes of one
// wavefront land in the same voxel block, and same-address LDS atomics
// serialise, so a small box is spread over copies that the dump sums.
// A copy is 16 B:  u64 (vx << 32 | density), u32 vy, u32 vz.
// Density addends are in [0, 1e6] and one window holds at most kW*(kL+1)
// points, so the low word never carries into vx: both sums are exact mod 2^32.
// ---------------------------------------------------------------------------
constexpr int kMaxCopies = 16;

// Where a lane's splat goes: its copy and the order it visits the corners of
// its 2x2x2 block in (lanes that share a copy take different orders).  Copy k of
// window voxel w is entry w * nc + k.  (Copy-major orders and separate
// planes per component were measured no faster.)
struct SplatLane {
    int slot, rot, nc;
};

// Copies of each voxel a window of `cells` voxels gets.
__device__ __forceinline__ SplatLane splat_lane(int cells)
{
    int nc = 1;
    while (nc < kMaxCopies && cells * nc * 2 <= (int)kWinCells) nc *= 2;
    const int lane = (int)threadIdx.x;
    return SplatLane{lane & (nc - 1), (lane / nc) & 7, nc};
}

// entry index of window voxel w in this lane's copy
__device__ __forceinline__ int splat_entry(int w, const SplatLane& sl) { return w * sl.nc + sl.slot; }

__device__ __forceinline__ void splat_cell(unsigned char* win, int e, float w, f3 v)
{
    const float SCALE = 1000000.0f;
    const f3 wv = smul(w, v);
    const uint32_t ix = (uint32_t)sat_i32(SCALE * wv.x);
    const uint32_t iy = (uint32_t)sat_i32(SCALE * wv.y);
    const uint32_t iz = (uint32_t)sat_i32(SCALE * wv.z);
    const uint32_t id = (uint32_t)sat_i32(SCALE * w);
    unsigned char* cellp = win + 16 * e;
    atomicAdd(reinterpret_cast<unsigned long long*>(cellp), ((unsigned long long)ix << 32) | id);
    atomicAdd(reinterpret_cast<uint32_t*>(cellp + 8), iy);
    atomicAdd(reinterpret_cast<uint32_t*>(cellp + 12), iz);
}

// Sum of window voxel w's copies as a GridCell (vx, vy, vz, density).
__device__ __forceinline__ int4 splat_sum(const int4* win, int w, const SplatLane& sl)
{
    uint64_t A = 0;
    uint32_t sy = 0, sz = 0;
    for (int k = 0; k < sl.nc; ++k) {
        const int4 raw = win[w * sl.nc + k];
        A += ((uint64_t)(uint32_t)raw.y << 32) | (uint32_t)raw.x;
        sy += (uint32_t)raw.z;
        sz += (uint32_t)raw.w;
    }
    return make_int4((int)(uint32_t)(A >> 32), (int)sy, (int)sz, (int)(uint32_t)A);
}

// Full 2x2x2 block inside the LDS window.
__device__ __forceinline__ void splat_block(int4* win, const Win& wd, const Rng& r, f3 q, f3 v, const SplatLane& sl)
{
    const float wx0 = wgt(q.x, r.x0), wx1 = wgt(q.x, r.x1);
    const float wy0 = wgt(q.y, r.y0), wy1 = wgt(q.y, r.y1);
    const float wz0 = wgt(q.z, r.z0), wz1 = wgt(q.z, r.z1);
    const int sx = sl.nc, sy = sx * wd.ex, sz = sy * wd.ey;
    const int e0 = splat_entry(win_index(wd, r.x0, r.y0, r.z0), sl);
    unsigned char* wb = reinterpret_cast<unsigned char*>(win);
    // Lanes that share a copy visit the corners in different orders (step i
    // takes corner i ^ rot), so lanes of one voxel block hit different cells
    // per instruction.  The lane's order is set up once: per axis the weight
    // and entry offset of step bit 0 and 1 are swapped when rot has that bit,
    // and every step then takes fixed registers.  Integer adds commute and
    // each corner's weight is still (wx * wy) * wz: the sums are unchanged.
    const bool fx = (sl.rot & 1) != 0, fy = (sl.rot & 2) != 0, fz = (sl.rot & 4) != 0;
    const float xa = fx ? wx1 : wx0, xb = fx ? wx0 : wx1;
    const float ya = fy ? wy1 : wy0, yb = fy ? wy0 : wy1;
    const float za = fz ? wz1 : wz0, zb = fz ? wz0 : wz1;
    const int ea = e0 + (fx ? sx : 0) + (fy ? sy : 0) + (fz ? sz : 0);  // step 0's corner
    const int dx = fx ? -sx : sx, dy = fy ? -sy : sy, dz = fz ? -sz : sz;
    const float wxy[4] = {xa * ya, xb * ya, xa * yb, xb * yb};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float w = wxy[i & 3] * ((i & 4) ? zb : za);
        splat_cell(wb, ea + ((i & 1) ? dx : 0) + ((i & 2) ? dy : 0) + ((i & 4) ? dz : 0), w, v);
    }
}

// Any non-empty block: into the LDS window when `lds`, else device atomics
// into the global overflow grid (when no window fits).
__device__ __forceinline__ void splat_any(int4* win, const Win& wd, bool lds, const SplatLane& sl,
                                          int4* __restrict__ g, const Rng& r, f3 q, f3 v, int G)
{
    const float SCALE = 1000000.0f;
#pragma unroll 1
    for (int a = r.x0; a <= r.x1; ++a) {
        const float wxa = wgt(q.x, a);
#pragma unroll 1
        for (int b = r.y0; b <= r.y1; ++b) {
            const float wxy = wxa * wgt(q.y, b);
#pragma unroll 1
            for (int c = r.z0; c <= r.z1; ++c) {
                const float w = wxy * wgt(q.z, c);
                if (lds) {
                    splat_cell(reinterpret_cast<unsigned char*>(win), splat_entry(win_index(wd, a, b, c), sl), w, v);
                } else {
                    const f3 wv = smul(w, v);
                    int* cell = reinterpret_cast<int*>(g + ((size_t)a + (size_t)b * G + (size_t)c * G * G));
                    atomicAdd(cell + 0, sat_i32(SCALE * wv.x));
                    atomicAdd(cell + 1, sat_i32(SCALE * wv.y));
                    atomicAdd(cell + 2, sat_i32(SCALE * wv.z));
                    atomicAdd(cell + 3, sat_i32(SCALE * w));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Force at a start-of-step position: gravity (+ wind, + strand noise) then the
// collider penalty loop (compute.comp:150-184).
// ---------------------------------------------------------------------------
// `mf`: the mesh pass's result for this point (used when COL has kColMesh)
template <int COL>
__device__ __forceinline__ f3 force_at(f3 x, f3 fs, const StepArgs& a, float4 mf)
{
    if (a.forcing) {
        // the shader's commented-out wind (compute.comp:150-151), after gravity
        float ft[3];
        forcing_term(Forcing{a.forcing, a.f_amp, a.f_a2, a.f_t3}, x.y, x.z, ft);
        fs = add(fs, mk(ft[0], ft[1], ft[2]));
    }
    f3 acc = mk(0.0f, 0.0f, 0.0f);
    int hits = 0;
    if (a.ncol > 0) {
        // sphere: compute.comp:160-169 (centre = transform[3].xyz)
        const f3 dv = sub(x, mk(a.sphere_c[0], a.sphere_c[1], a.sphere_c[2]));
        const float d2 = dot(dv, dv);
        // d2 < sphere_r2 is exactly sqrt(d2) < sphere_r
        if (d2 < a.sphere_r2) {
            const float dist = sqrtf(d2);
            const float d = a.sphere_r - dist;
            const f3 n = muls(dv, 1.0f / dist);
            acc = add(acc, smul(a.k * d, n));
            hits = 1;
        }
    }
    if (COL & kColEll) {
        // ellipsoids: EllipsoidCollision / GetPointOnEllipsoid /
        // GetEllipsoidNormal (compute.comp:64-79, 170-179)
#pragma unroll 1
        for (uint32_t j = 1; j < a.ncol; ++j) {
            const float* T = a.cols[j].m;
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
    if (COL & kColMesh) {
        // the mesh counts as one more collider, after the reference's ones
        // (its contact computed by the mesh pass)
        if (mf.w != 0.0f) {
            acc = add(acc, mk(mf.x, mf.y, mf.z));
            hits += 1;
        }
    }
    f3 force = fs;
    if (hits > 0) {
        const float m = (float)hits;
        force = add(force, mk(acc.x / m, acc.y / m, acc.z / m));
    }
    return force;
}

// Pointers a kernel reads out of device memory (the batched launches' argument
// blob) are generic to the compiler, which then emits flat accesses: they wait
// on the LDS counter as well as the memory counter, so every LDS wait in the
// step body would also wait for outstanding HBM loads.  Every buffer here is
// global memory: reading each pointer field as a global-address-space pointer
// lets the compiler emit global loads, stores and atomics through it, as it
// does for pointers taken from kernel arguments.
template <class T>
__device__ __forceinline__ T* global_ptr(T* const& field)
{
    using GT = __attribute__((address_space(1))) T*;
    return (T*)*static_cast<const GT*>(static_cast<const void*>(&field));
}

__device__ __forceinline__ GridWork global_grid(const GridWork& in)
{
    GridWork w = in;
    w.grid_i = global_ptr(in.grid_i);
    w.grid_f = global_ptr(in.grid_f);
    w.pool = global_ptr(in.pool);
    w.desc = global_ptr(in.desc);
    w.counters = global_ptr(in.counters);
    w.overflow = global_ptr(in.overflow);
    w.gbox = global_ptr(in.gbox);
    w.brick_cnt = global_ptr(in.brick_cnt);
    w.brick_list = global_ptr(in.brick_list);
    w.brick_nz = global_ptr(in.brick_nz);
    w.near_limit = global_ptr(in.near_limit);
    w.dlist = global_ptr(in.dlist);
    w.dclaim = global_ptr(in.dclaim);
    w.dbox = global_ptr(in.dbox);
    return w;
}

__device__ __forceinline__ MeshArgs global_mesh(const MeshArgs& in)
{
    MeshArgs m = in;
    m.tris = global_ptr(in.tris);
    m.cell_start = global_ptr(in.cell_start);
    m.cell_recs = global_ptr(in.cell_recs);
    return m;
}

// A groom's arguments read from the blob, every pointer global.
__device__ __forceinline__ StepArgs global_args(const StepArgs& in)
{
    StepArgs a = in;
    a.pos = global_ptr(in.pos);
    a.vel = global_ptr(in.vel);
    a.corr = global_ptr(in.corr);
    a.seg = global_ptr(in.seg);
    a.noise = global_ptr(in.noise);
    a.gw = global_grid(in.gw);
    a.cols = global_ptr(in.cols);
    a.mesh = global_mesh(in.mesh);
    a.mesh_f = global_ptr(in.mesh_f);
    return a;
}

// Many grooms in one launch: wg_start[g] is the first global workgroup of
// groom g (prefix sum of their workgroup counts, ngroom+1 entries).
__device__ __forceinline__ uint32_t find_groom(const uint32_t* __restrict__ start, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;  // start[lo] <= x < start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (start[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

}  // namespace hs
