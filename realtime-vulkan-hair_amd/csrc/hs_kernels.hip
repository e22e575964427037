// hs_kernels.hip - HIP kernels of the strand-physics pass for gfx950.
//
// State layout in HBM (strand-major SoA, one plane per component):
//   pos[c][i][p], vel[c][i][p], c in {x,y,z}, i = point 0..N-1, p = physical
//   strand slot 0..Sp-1.  Strands are stored in Morton order of their roots
//   (perm[p] = caller's strand index), so the kW strands of one workgroup are
//   spatial neighbours.  At a fixed point index a wavefront reads 64
//   consecutive floats of each plane: one 256 B line per plane.
//
// One step = two kernels.
//  k_step   (one workgroup per kW strands) walks each strand root to tip in
//           segments of kL points held in registers.  Per segment:
//             P: per point, friction with the previous step's voxel window
//                staged in LDS (compute.comp:257-298), force and predicted
//                position (compute.comp:150-187), all points independent;
//             F: follow-the-leader, damped velocity, DFTL (compute.comp:
//                190-215), the root-to-tip chain;
//             S: trilinear fixed-point splat (compute.comp:217-252) of the
//                points whose velocity is final into an LDS window, dumped as
//                a dense partial box to a pool.
//  k_reduce (one workgroup per 8^3 voxel brick) sums every partial box that
//           overlaps its brick and writes the step's grid: the GridCell ints
//           and the float cells the next step's G phase reads.  It writes every
//           voxel, which also clears the grid (the vkCmdFillBuffer of
//           Renderer.cpp:2063).
// The step is skewed: the friction of step k is applied by k_step of step
// k+1 (or by the read-back kernel).  Per point the operations and their order
// are those of the unskewed shader, so results are bit-identical; integer
// voxel sums are order-independent (wrapping adds), so the grid is exact too.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device_math.hpp"
#include "hs_forcing.hpp"
#include "hs_kernels.h"

namespace hs {

static_assert((uint64_t)kW * (kL + 1) * 1000000ull < (1ull << 32),
              "packed (vx<<32 | density) LDS sums need the density of one window below 2^32");
static_assert(kW % 64 == 0, "workgroup is whole wavefronts");
static_assert(kL <= 31, "per-point flags live in one 32-bit mask");

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

__device__ __forceinline__ bool rng_empty(const Rng& r) { return r.x0 > r.x1 || r.y0 > r.y1 || r.z0 > r.z1; }

__device__ __forceinline__ bool rng_full(const Rng& r)
{
    return r.x1 == r.x0 + 1 && r.y1 == r.y0 + 1 && r.z1 == r.z0 + 1;
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
    return r.x0 >= w.x0 && r.x1 < w.x0 + w.ex && r.y0 >= w.y0 && r.y1 < w.y0 + w.ey && r.z0 >= w.z0 &&
           r.z1 < w.z0 + w.ez;
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
// segments of the longest strand (256 points): the most deferred-list
// entries one workgroup can publish in a step
#define HS_MAX_OWN_DEFERRED ((256 - 1 + HS_TILE_POINTS - 1) / HS_TILE_POINTS)

// Where a lane's splat goes: its copy and the corner it starts its 2x2x2
// block at (lanes that share a copy start at different corners).  Copy k of
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
    // Lanes that share a copy start at different corners, so lanes of one
    // voxel block hit different cells per instruction.  Integer adds commute
    // and each corner's weight is still (wx * wy) * wz: the sums are unchanged.
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = (i + sl.rot) & 7;
        const bool dx = (k & 1) != 0, dy = (k & 2) != 0, dz = (k & 4) != 0;
        const float w = ((dx ? wx1 : wx0) * (dy ? wy1 : wy0)) * (dz ? wz1 : wz0);
        splat_cell(wb, e0 + (dx ? sx : 0) + (dy ? sy : 0) + (dz ? sz : 0), w, v);
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
// Closest point on triangle abc to x (Ericson, Real-Time Collision Detection,
// 2005, 5.1.5) and the feature it lies on: 0..2 vertex a/b/c, 3 edge ab,
// 4 edge bc, 5 edge ca, 6 face.
__device__ __forceinline__ f3 closest_on_tri(f3 x, f3 pa, f3 pb, f3 pc, int* feature)
{
    const f3 ab = sub(pb, pa), ac = sub(pc, pa), ap = sub(x, pa);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        *feature = 0;
        return pa;
    }
    const f3 bp = sub(x, pb);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        *feature = 1;
        return pb;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        *feature = 3;
        return add(pa, smul(d1 / (d1 - d3), ab));
    }
    const f3 cp = sub(x, pc);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        *feature = 2;
        return pc;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        *feature = 5;
        return add(pa, smul(d2 / (d2 - d6), ac));
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        *feature = 4;
        return add(pb, smul((d4 - d3) / ((d4 - d3) + (d5 - d6)), sub(pc, pb)));
    }
    const float denom = 1.0f / ((va + vb) + vc);
    const float v = vb * denom, w = vc * denom;
    *feature = 6;
    return add(add(pa, smul(v, ab)), smul(w, ac));
}

__device__ __forceinline__ f3 ld_f3(const float4* p)
{
    const float4 v = *p;
    return mk(v.x, v.y, v.z);
}

// Mesh contact: the mesh point q nearest to x among the triangles listed for
// x's grid cell (every triangle within the margin of the cell is listed, so
// within the margin this is the nearest over the whole mesh; lowest id wins
// ties); contact when |x - q|^2 < margin^2 and dot(x - q, n) < 0 for the
// pseudo-normal n of q's feature (Baerentzen & Aanaes, IEEE TVCG 2005).
// Returns (k |x - q|) normalize(q - x) in xyz and 1 in w on contact, else 0.
// No reference counterpart.  Evaluated by the mesh pass (k_mesh_force /
// k_mesh_force_batch) for every stored point before the step kernel, which
// reads the result.
#ifndef HS_MESH_PASS_ROUND
#define HS_MESH_PASS_ROUND 2
#endif
constexpr uint32_t kMeshRound = HS_MESH_PASS_ROUND;  // mesh candidates loaded per round

__device__ __forceinline__ float4 mesh_contact(const MeshArgs& M, float px, float py, float pz, float k)
{
    const f3 x = mk(px, py, pz);
    const float4 none = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float gx = (x.x - M.ox) * M.icx, gy = (x.y - M.oy) * M.icy, gz = (x.z - M.oz) * M.icz;
    // outside the grid (or NaN): farther than the margin from every triangle
    if (!(gx >= 0.0f && gx < (float)M.dx && gy >= 0.0f && gy < (float)M.dy && gz >= 0.0f && gz < (float)M.dz))
        return none;
    const uint32_t cx = min((uint32_t)gx, M.dx - 1), cy = min((uint32_t)gy, M.dy - 1),
                   cz = min((uint32_t)gz, M.dz - 1);
    const uint32_t cell = (cz * M.dy + cy) * M.dx + cx;
    const uint32_t i0 = M.cell_start[cell], i1 = M.cell_start[cell + 1];
    // distance from x to its cell's centre: a candidate's distance is at least
    // its key minus this (the key bounds the distance from the centre)
    const f3 cc = mk(M.ox + ((float)cx + 0.5f) / M.icx, M.oy + ((float)cy + 0.5f) / M.icy,
                     M.oz + ((float)cz + 0.5f) / M.icz);
    const float dxc = length(sub(x, cc));
    float best = M.margin2;
    // reject radius: a triangle whose bounding sphere lies farther than
    // sqrt(best) plus a slack (1e-5 (1 + |x|) absolute and 1e-4 relative on
    // the squares, far above the float rounding of a point-triangle distance
    // at these magnitudes) cannot be strictly nearer than `best`
    const float slack = 1e-5f * (1.0f + fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fabsf(x.z)));
    float reach = sqrtf(best) + slack;
    int best_feat = 0;
    uint32_t best_t = 0xffffffffu;
    f3 best_q = mk(0.0f, 0.0f, 0.0f);
    // candidates kMeshRound at a time: their ids, spheres and keys are contiguous in
    // the cell's list, so the loads are in flight together.  The list is
    // sorted by `key` (the distance from the cell's centre to the triangle's
    // bounding sphere, rounded down); key - |x - centre| bounds the distance
    // from below, so once that passes `reach` no later entry can win.  The
    // nearest of equal distances is the lowest id (the order of the list
    // does not matter).
    bool more = true;
    for (uint32_t i = i0; i < i1 && more; i += kMeshRound) {
        float4 sp[kMeshRound];
        uint32_t tt[kMeshRound];
        float kk[kMeshRound];
#pragma unroll
        for (uint32_t u = 0; u < kMeshRound; ++u) {
            const uint32_t e = min(i + u, i1 - 1);
            sp[u] = M.cell_spheres[e];
            tt[u] = M.cell_tris[e];
            kk[u] = M.cell_keys[e];
        }
#pragma unroll
        for (uint32_t u = 0; u < kMeshRound; ++u) {
            if (i + u >= i1 || !more) break;
            if (kk[u] - dxc > reach) {
                more = false;
                break;
            }
            const f3 dc = sub(x, mk(sp[u].x, sp[u].y, sp[u].z));
            const float lim = sp[u].w + reach;
            if (dot(dc, dc) > lim * lim * 1.0001f + 1e-12f) continue;
            const float4* r = M.tris + (size_t)tt[u] * 10;
            int feat;
            const f3 q = closest_on_tri(x, ld_f3(r), ld_f3(r + 1), ld_f3(r + 2), &feat);
            const f3 dq = sub(x, q);
            const float d2 = dot(dq, dq);
            if (d2 < best || (d2 == best && best_t != 0xffffffffu && tt[u] < best_t)) {
                best = d2;
                reach = sqrtf(best) + slack;
                best_t = tt[u];
                best_feat = feat;
                best_q = q;
            }
        }
    }
    if (best_t == 0xffffffffu) return none;
    // feature -> record slot: vertices 7..9, edges 4..6, face 3
    const int slot = best_feat < 3 ? 7 + best_feat : (best_feat < 6 ? best_feat + 1 : 3);
    const f3 n = ld_f3(M.tris + (size_t)best_t * 10 + slot);
    if (!(dot(sub(x, best_q), n) < 0.0f)) return none;
    const float d = sqrtf(best);
    const f3 c = smul(k * d, normalize(sub(best_q, x)));
    return make_float4(c.x, c.y, c.z, 1.0f);
}

// Mesh pass: the mesh contact of every stored point (the positions a step
// starts from), one thread per point, into out[N][Sp].  Apart from the step
// kernel, whose register budget holds it to two waves per SIMD, this runs at
// full occupancy, which hides the scattered triangle loads.
__device__ __forceinline__ void mesh_point(const MeshArgs& M, const float* __restrict__ pos, uint32_t plane, float k,
                                           float4* __restrict__ out, uint32_t i)
{
    out[i] = mesh_contact(M, pos[i], pos[plane + i], pos[2 * (size_t)plane + i], k);
}

__global__ __launch_bounds__(256) void k_mesh_force(MeshArgs M, const float* __restrict__ pos, uint32_t plane,
                                                    float k, float4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < plane) mesh_point(M, pos, plane, k, out, i);
}

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

// ---------------------------------------------------------------------------
// Workgroup reductions and window planning
// ---------------------------------------------------------------------------
constexpr int kWaves = kW / 64;

// Minimum over the wavefront (every lane active): DPP steps within each row
// of 16 lanes (VALU moves, no LDS traffic), then the four row results.
#ifndef HS_DPP_MIN
#define HS_DPP_MIN 1
#endif
__device__ __forceinline__ int wave_min(int v)
{
#if !HS_DPP_MIN
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
    return v;
#endif
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Box of voxels (inclusive bounds).  Empty: lo = INT_MAX, hi = INT_MIN.
struct Box {
    int lo[3], hi[3];
};

__device__ __forceinline__ void box_add(Box& b, const Rng& r)
{
    if (rng_empty(r)) return;
    b.lo[0] = min(b.lo[0], r.x0);
    b.hi[0] = max(b.hi[0], r.x1);
    b.lo[1] = min(b.lo[1], r.y0);
    b.hi[1] = max(b.hi[1], r.y1);
    b.lo[2] = min(b.lo[2], r.z0);
    b.hi[2] = max(b.hi[2], r.z1);
}

__device__ __forceinline__ Box box_empty()
{
    Box b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = INT32_MAX;
        b.hi[k] = INT32_MIN;
    }
    return b;
}

__device__ __forceinline__ bool box_is_empty(const Box& b)
{
    return b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2];
}

// Reduce a box over the workgroup; every thread gets the result.
__device__ __forceinline__ Box block_box(Box b, int* scratch)
{
    int v[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = b.lo[k];
        v[k + 3] = b.hi[k] == INT32_MIN ? INT32_MAX : -b.hi[k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = wave_min(v[k]);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) scratch[wave * 6 + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int m = scratch[k];
        for (int w = 1; w < kWaves; ++w) m = min(m, scratch[w * 6 + k]);
        v[k] = m;
    }
    __syncthreads();
    Box r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = v[k];
        r.hi[k] = v[k + 3] == INT32_MAX ? INT32_MIN : -v[k + 3];
    }
    return r;
}

// Chunking of a box into LDS windows of at most `cap` voxels: full x extent,
// chunks in y and z overlapping by one voxel so that every 2x2x2 block of the
// box lies whole inside at least one chunk.
struct Plan {
    int ex, ey, ez;   // box extents
    int cy, cz;       // chunk extents in y and z (cells)
    int ny, nz;       // chunk counts
    bool ok;          // false: no chunking fits (or empty box)
};

__device__ __forceinline__ Plan plan_box(const Box& b, int cap, bool splat_plan = false)
{
    Plan p;
    if (box_is_empty(b)) {
        p.ex = p.ey = p.ez = 0;
        p.cy = p.cz = 2;
        p.ny = p.nz = 0;
        p.ok = false;
        return p;
    }
    p.ex = b.hi[0] - b.lo[0] + 1;
    p.ey = b.hi[1] - b.lo[1] + 1;
    p.ez = b.hi[2] - b.lo[2] + 1;
    p.ok = true;
    if (p.ex * p.ey * p.ez <= cap) {
        p.cy = p.ey;
        p.cz = p.ez;
#ifdef HS_NO_CHUNKS
    } else if (splat_plan) {
        p.ok = false;
        p.cy = p.cz = 2;
#endif
    } else if (p.ex * p.ey * 2 <= cap) {
        p.cy = p.ey;
        p.cz = cap / (p.ex * p.ey);
    } else if (p.ex * 4 <= cap) {
        p.cz = 2;
        p.cy = cap / (p.ex * 2);
    } else {
        p.ok = false;
        p.cy = p.cz = 2;
    }
    p.ny = p.cy >= p.ey ? 1 : 1 + (p.ey - 2) / (p.cy - 1);
    p.nz = p.cz >= p.ez ? 1 : 1 + (p.ez - 2) / (p.cz - 1);
    return p;
}

__device__ __forceinline__ Win chunk_of(const Box& b, const Plan& p, int iy, int iz)
{
    Win w;
    w.x0 = b.lo[0];
    w.ex = p.ex;
    w.y0 = b.lo[1] + iy * (p.cy - 1);
    w.z0 = b.lo[2] + iz * (p.cz - 1);
    w.ey = min(p.cy, b.hi[1] - w.y0 + 1);
    w.ez = min(p.cz, b.hi[2] - w.z0 + 1);
    return w;
}

// ---------------------------------------------------------------------------
// Partial boxes: an LDS window is summed over its copies and dumped into the
// pool, and its box is listed with every reduce brick it overlaps.
// ---------------------------------------------------------------------------
// `own` >= 0: the dump goes to that workgroup-segment's own pool slot and
// descriptor (no allocator); otherwise to the shared part, taken with atomics.
// Every thread of the workgroup calls this after the window is complete.
__device__ __forceinline__ void dump_window(const StepArgs& a, const Win& cw, int cells, const SplatLane& sl, int own, int4* win,
                            uint32_t* bcast)
{
    uint32_t* ctr = a.gw.counters + a.parity * 4;
    if (threadIdx.x == 0) {
        uint32_t off, di;
        if (own >= 0) {
            off = (uint32_t)own * kSlotCells;
            di = (uint32_t)own;
        } else {
            off = a.gw.nws * kSlotCells + atomicAdd(ctr + 0, (uint32_t)cells);
            di = a.gw.nws + atomicAdd(ctr + 1, 1u);
        }
        const bool fits = (uint64_t)off + cells <= a.gw.pool_cap && di < a.gw.desc_cap;
        if (di < a.gw.desc_cap) {
            int* d = a.gw.desc + ((size_t)a.parity * a.gw.desc_cap + di) * kDescInts;
            d[0] = cw.x0;
            d[1] = cw.y0;
            d[2] = cw.z0;
            d[3] = fits ? cw.ex : 0;  // ex == 0: empty descriptor
            d[4] = cw.ey;
            d[5] = cw.ez;
            d[6] = (int)off;
            d[7] = 0;
        }
        if (!fits) atomicOr(ctr + 2, 1u);
        bcast[0] = off;
        bcast[1] = fits ? 1u : 0u;
        bcast[2] = di;
    }
    __syncthreads();
    const uint32_t off = bcast[0];
    const bool fits = bcast[1] != 0;
    const int G = (int)a.G;
    if (fits) {
        // list the box with every brick it overlaps (one device atomic each, in parallel)
        const uint32_t di = bcast[2];
        const int bx0 = cw.x0 / (int)kBX, by0 = cw.y0 / (int)kBY, bz0 = cw.z0 / (int)kBZ;
        const int nbx = (cw.x0 + cw.ex - 1) / (int)kBX - bx0 + 1;
        const int nby = (cw.y0 + cw.ey - 1) / (int)kBY - by0 + 1;
        const int nbz = (cw.z0 + cw.ez - 1) / (int)kBZ - bz0 + 1;
        const uint32_t nbr = a.gw.nbx * a.gw.nby * a.gw.nbz;
        for (int t = threadIdx.x; t < nbx * nby * nbz; t += kW) {
            const uint32_t b = ((uint32_t)(bz0 + t / (nbx * nby)) * a.gw.nby + (uint32_t)(by0 + (t / nbx) % nby)) *
                                   a.gw.nbx +
                               (uint32_t)(bx0 + t % nbx);
            const uint32_t at = atomicAdd(a.gw.brick_cnt + a.parity * nbr + b, 1u);
            if (at < kBrickList) a.gw.brick_list[((size_t)a.parity * nbr + b) * kBrickList + at] = di;
        }
    }
    int4* dst = a.gw.pool + (size_t)a.parity * a.gw.pool_cap + off;
    for (int c = threadIdx.x; c < cells; c += kW) {
        const int4 cell = splat_sum(win, c, sl);
        if (fits) {
            dst[c] = cell;
        } else if (cell.x | cell.y | cell.z | cell.w) {
            // no room in the pool: device atomics into the overflow grid
            const int ax = c % cw.ex;
            const int rest = c / cw.ex;
            const size_t gi = (size_t)(cw.x0 + ax) + (size_t)(cw.y0 + rest % cw.ey) * G +
                              (size_t)(cw.z0 + rest / cw.ey) * G * G;
            int* gcell = reinterpret_cast<int*>(a.gw.overflow + gi);
            atomicAdd(gcell + 0, cell.x);
            atomicAdd(gcell + 1, cell.y);
            atomicAdd(gcell + 2, cell.z);
            atomicAdd(gcell + 3, cell.w);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Deferred chunks.  A workgroup-segment whose voxel box needs several windows
// records its box and chunking (dbox[ws]), then appends ws to the step's list
// (dlist[parity][k] = step tag, chunk count, ws; counters[parity][3] entries).  Chunk c of
// list entry k is taken by whoever first raises dclaim[parity][k] past c, and
// splatted from the segment's stored positions and velocities.  Each point
// goes to exactly one chunk: the one whose y/z range starts at or before its
// block and reaches past it (chunk starts step by window size - 1, so that is
// the quotient below).  Integer voxel sums commute, so the grid equals the
// one the segment would have splatted itself.
// Every workgroup of the groom, when done with its own strands, reads the
// list once and takes any chunk still free; the list's owners additionally
// take all chunks of their own entries that are left, so every chunk is done
// before the last owner finishes, without any workgroup waiting on another.
// `tag` (a per-groom step count) tells this step's entries from older ones
// that share the slot.
// ---------------------------------------------------------------------------
// A list entry: step tag (28 bits) | chunk count (12 bits) | workgroup-segment (24 bits).
constexpr uint32_t kMaxDeferredChunks = 4095;
__device__ __forceinline__ unsigned long long dentry(uint32_t tag, uint32_t n, uint32_t ws)
{
    return ((unsigned long long)(tag & 0x0fffffffu) << 36) | ((unsigned long long)n << 24) | ws;
}
__device__ __forceinline__ bool dentry_is(unsigned long long e, uint32_t tag) { return (uint32_t)(e >> 36) == (tag & 0x0fffffffu); }
__device__ __forceinline__ uint32_t dentry_n(unsigned long long e) { return (uint32_t)(e >> 24) & 0xfffu; }
__device__ __forceinline__ uint32_t dentry_ws(unsigned long long e) { return (uint32_t)e & 0xffffffu; }

__device__ __forceinline__ uint32_t publish_deferred(const StepArgs& a, uint32_t ws, const Box& box, const Plan& pl)
{
    int* db = a.gw.dbox + (size_t)ws * kDBoxInts;
    for (int k = 0; k < 3; ++k) {
        db[k] = box.lo[k];
        db[3 + k] = box.hi[k];
    }
    db[6] = pl.cy;
    db[7] = pl.cz;
    db[8] = pl.ny;
    db[9] = pl.nz;
    __threadfence();  // the plan (and the segment's stored state) before the list entry
    uint32_t* ctr = a.gw.counters + a.parity * 4;
    const uint32_t k = atomicAdd(ctr + 3, 1u);
    __hip_atomic_store(a.gw.dlist + (size_t)a.parity * a.gw.nws + k, dentry(a.tag, (uint32_t)(pl.ny * pl.nz), ws),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return k;
}

// Chunk c of the deferred workgroup-segment ws (plan at db): zero the window,
// splat the segment's finished points whose blocks fall in this chunk, dump.
__device__ __forceinline__ void splat_deferred_chunk(const StepArgs& a, uint32_t ws, const int* db, int c, int4* win, uint32_t* bcast)
{
    Box box;
    for (int k = 0; k < 3; ++k) {
        box.lo[k] = db[k];
        box.hi[k] = db[3 + k];
    }
    Plan pl;
    pl.ex = box.hi[0] - box.lo[0] + 1;
    pl.ey = box.hi[1] - box.lo[1] + 1;
    pl.ez = box.hi[2] - box.lo[2] + 1;
    pl.cy = db[6];
    pl.cz = db[7];
    pl.ny = db[8];
    pl.nz = db[9];
    pl.ok = true;
    const int iy = c % pl.ny, iz = c / pl.ny;
    const Win cw = chunk_of(box, pl, iy, iz);
    const int cells = cw.ex * cw.ey * cw.ez;
    const SplatLane sl = splat_lane(cells);
    for (int e = threadIdx.x; e < cells * sl.nc; e += kW) win[e] = make_int4(0, 0, 0, 0);
    __syncthreads();
    const uint32_t wgo = ws / a.nseg, sg = ws % a.nseg;
    const uint32_t s = wgo * kW + threadIdx.x;
    const uint32_t N = a.N, Sp = a.Sp;
    const uint32_t f = 1 + sg * kL;
    const uint32_t cnt = min(kL, N - f);
    const bool last_seg = f + cnt == N;
    if (s < a.S) {
        const uint32_t P4 = N * Sp * 4u, P8 = N * Sp * 8u;
        const f3 origin = mk(a.ox, a.oy, a.oz);
        // finished points of the segment: f-1 (carried), f..f+cnt-2, and the
        // tip; four at a time, their loads in flight together
        const uint32_t i0 = f >= 2 ? f - 1 : f;
        const uint32_t i1 = last_seg ? f + cnt - 1 : f + cnt - 2;
        constexpr uint32_t kGrp = 4;
#pragma unroll 1
        for (uint32_t g0 = i0; g0 <= i1; g0 += kGrp) {
        f3 XP[kGrp], VP[kGrp];
#pragma unroll
        for (uint32_t t = 0; t < kGrp; ++t) {
            if (g0 + t > i1) continue;
            const uint32_t bo = ((g0 + t) * Sp + s) * 4u;
            XP[t] = ld3(a.pos, bo, P4, P8);
            VP[t] = ld3(a.vel, bo, P4, P8);
        }
#pragma unroll
        for (uint32_t t = 0; t < kGrp; ++t) {
            if (g0 + t > i1) continue;
            const f3 x = XP[t];
            const f3 v = VP[t];
            const f3 q = grid_coord(x, origin, a.inv_h);
            const Rng rg = point_range(q, (int)a.G);
            if (rng_empty(rg)) continue;
            const int jy = pl.ny > 1 ? min(pl.ny - 1, (rg.y0 - box.lo[1]) / (pl.cy - 1)) : 0;
            const int jz = pl.nz > 1 ? min(pl.nz - 1, (rg.z0 - box.lo[2]) / (pl.cz - 1)) : 0;
            if (jy != iy || jz != iz) continue;
            if (rng_full(rg))
                splat_block(win, cw, rg, q, v, sl);
            else
                splat_any(win, cw, true, sl, a.gw.overflow, rg, q, v, (int)a.G);
        }
        }
    }
    __syncthreads();
    dump_window(a, cw, cells, sl, -1, win, bcast);
}

// Run by every workgroup at its end (all threads): help with, or (for its own
// entries) finish, the step's deferred chunks.  The list is read with one
// thread per entry, so a workgroup with nothing to take is done after one
// round trip.  Workgroups finish at about the same time, so a helper starts
// at an entry picked by its index, takes at most one chunk and gives up after
// two entries with none left; the owners then take every chunk of their own
// entries that is still free.
__device__ __forceinline__ bool take_chunk(const StepArgs& a, uint32_t kk, int4* win, uint32_t* bcast)
{
    const unsigned long long* list = a.gw.dlist + (size_t)a.parity * a.gw.nws;
    uint32_t* claim = a.gw.dclaim + (size_t)a.parity * a.gw.nws;
    if (threadIdx.x == 0) {
        const unsigned long long e = __hip_atomic_load(list + kk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bcast[0] = dentry_ws(e);
        bcast[1] = atomicAdd(claim + kk, 1u);
        bcast[2] = dentry_n(e);
        __threadfence();  // the plan and the owner's stored points after its list entry
    }
    __syncthreads();
    const uint32_t ws = bcast[0], c = bcast[1], n = bcast[2];
    __syncthreads();
    if (c >= n) return false;
    splat_deferred_chunk(a, ws, a.gw.dbox + (size_t)ws * kDBoxInts, (int)c, win, bcast);
    return true;
}

__device__ __forceinline__ void deferred_chunks(const StepArgs& a, uint32_t wg, int4* win, uint32_t* bcast,
                                                uint32_t* want, const uint32_t* own_k, const uint32_t* n_own_p)
{
    uint32_t* ctr = a.gw.counters + a.parity * 4;
    __syncthreads();
    if (threadIdx.x == 0) bcast[0] = __hip_atomic_load(ctr + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t nd = bcast[0];
    const uint32_t n_own = *n_own_p;  // written by thread 0 before the barrier above
    if (nd == 0) return;
    const unsigned long long* list = a.gw.dlist + (size_t)a.parity * a.gw.nws;
    const uint32_t* claim = a.gw.dclaim + (size_t)a.parity * a.gw.nws;
    // want[0]: entries of other workgroups with chunks left, their indices in
    // want[1..]; at most kW - 1 per pass
    for (uint32_t base = 0; base < nd; base += kW - 1) {
        if (threadIdx.x == 0) want[0] = 0;
        __syncthreads();
        const uint32_t k = base + threadIdx.x;
        if (threadIdx.x < kW - 1 && k < nd) {
            const unsigned long long e = __hip_atomic_load(list + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t taken = __hip_atomic_load(claim + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (dentry_is(e, a.tag) && dentry_ws(e) / a.nseg != wg && taken < dentry_n(e))
                want[1 + atomicAdd(want, 1u)] = k;
        }
        __syncthreads();
        const uint32_t nw = want[0];
        // help with one chunk: try two entries, starting at one picked by this
        // workgroup's index
        for (uint32_t tried = 0, j = wg % max(nw, 1u); tried < min(nw, 2u); ++tried, j = (j + 1) % nw)
            if (take_chunk(a, want[1 + j], win, bcast)) break;
        __syncthreads();
    }
    // its own entries (recorded when published): everything left
    for (uint32_t j = 0; j < n_own; ++j)
        while (take_chunk(a, own_k[j], win, bcast)) {
        }
}

// ---------------------------------------------------------------------------
// The step kernel
// ---------------------------------------------------------------------------
// XCD-aware order: dispatch places consecutive blockIdx on different XCDs
// (round robin over 8); give each XCD a contiguous run of logical workgroups
// (Morton-ordered strand blocks) so neighbouring windows share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t block, uint32_t nblocks)
{
    const uint32_t per = (nblocks + 7) / 8;
    return (block % 8) * per + block / 8;
}

// One workgroup's share of a step: strands [wg*kW, wg*kW + kW) of the groom `a`.
template <bool GATHER, bool SPLAT, int COL>
__device__ __forceinline__ void step_body(const StepArgs& a, uint32_t wg)
{
    __shared__ __align__(16) int4 win[kWinCells];
    __shared__ int red[kW > kWaves * 6 ? kW : kWaves * 6];  // box reductions; the deferred-list scan
    __shared__ uint32_t bcast[3];
    __shared__ uint32_t own_k[HS_MAX_OWN_DEFERRED];  // deferred-list entries this workgroup published
    __shared__ uint32_t n_own[1];
    if (SPLAT && threadIdx.x == 0) n_own[0] = 0;

    const uint32_t s = wg * kW + threadIdx.x;
    const bool active = s < a.S;
    const uint32_t Sp = a.Sp;
    const uint32_t N = a.N;
    const uint32_t plane = N * Sp;
    // byte offsets into the [3][N][Sp] arrays stay below 2^32 (checked on the
    // host), so every access is an SGPR base plus a 32-bit VGPR offset
    const uint32_t P4 = plane * 4u, P8 = plane * 8u;
    float* __restrict__ posb = a.pos;
    float* __restrict__ velb = a.vel;
    const f3 origin = mk(a.ox, a.oy, a.oz);
    const int G = (int)a.G;

    f3 par = mk(0.0f, 0.0f, 0.0f);
    float r = 0.0f;
    f3 fs = mk(a.fx, a.fy, a.fz);
    if (active) {
        par = ld3(posb, s * 4u, P4, P8);
        r = a.seg[s];
        if (a.noise) fs = add(fs, smul(a.noise[s], mk(a.ndx, a.ndy, a.ndz)));
    }
    // point f-1 of the previous segment: position final, velocity awaiting DFTL
    f3 cp = mk(0.0f, 0.0f, 0.0f), cv = mk(0.0f, 0.0f, 0.0f);

    for (uint32_t sg = 0; sg < a.nseg; ++sg) {
        const uint32_t f = 1 + sg * kL;
        const uint32_t cnt = min(kL, N - f);
        const bool last_seg = f + cnt == N;
        const uint32_t b0 = (f * Sp + s) * 4u;  // byte offset of point f
        const uint32_t bS = Sp * 4u;             // byte stride between points
        f3 X[kL], U[kL];
#pragma unroll
        for (uint32_t t = 0; t < kL; ++t) {
            X[t] = mk(0.0f, 0.0f, 0.0f);
            U[t] = mk(0.0f, 0.0f, 0.0f);
            if (active && t < cnt) {
                X[t] = ld3(posb, b0 + t * bS, P4, P8);
                U[t] = ld3(velb, b0 + t * bS, P4, P8);
            }
        }

        // ---------------- G: stage the previous step's voxel window ----------------
        Win gwin{0, 0, 0, 0, 0, 0};
        bool gwin_ok = false;
        if (GATHER) {
            const int* gb = a.gw.gbox + ((size_t)wg * a.nseg + sg) * 6;
            Box box;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                box.lo[k] = gb[k];
                box.hi[k] = gb[3 + k];
            }
            const Plan pl = plan_box(box, (int)min(a.win_gather, kWinCells));
            if (pl.ok && pl.ny == 1 && pl.nz == 1) {
                gwin = chunk_of(box, pl, 0, 0);
                gwin_ok = true;
                const int cells = gwin.ex * gwin.ey * gwin.ez;
                float4* fw = reinterpret_cast<float4*>(win);
                // loads for up to 8 cells per thread in flight before the LDS stores
                const uint32_t GG = (uint32_t)G * (uint32_t)G;
                for (int c0 = 0; c0 < cells; c0 += 8 * kW) {
                    // unconditional (clamped) loads: all 8 in flight before the stores
                    float4 tmp[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const uint32_t c = (uint32_t)min(c0 + k * (int)kW + (int)threadIdx.x, cells - 1);
                        const uint32_t ax = c % (uint32_t)gwin.ex;
                        const uint32_t rest = c / (uint32_t)gwin.ex;
                        const uint32_t gi = (uint32_t)gwin.x0 + ax + ((uint32_t)gwin.y0 + rest % (uint32_t)gwin.ey) * (uint32_t)G +
                                            ((uint32_t)gwin.z0 + rest / (uint32_t)gwin.ey) * GG;
                        tmp[k] = a.gw.grid_f[gi];
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int c = c0 + k * kW + threadIdx.x;
                        if (c < cells) fw[c] = tmp[k];
                    }
                }
                __syncthreads();
            }
        }

        // ------- P: friction, force and predicted position of every point -------
        // Previous step's phase 3 (compute.comp:257-298) for each point: from
        // the LDS window, or (a workgroup whose box exceeds any window)
        // straight from the float grid, the 8 cells of a block loaded together.
        auto predict = [&](uint32_t t, f3 gv) {
            const f3 x = X[t];
            f3 u = U[t];
            if (GATHER) u = friction(u, gv, a.mu_prev, a.omu_prev);
            // predict (compute.comp:187)
            float4 mf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (COL & kColMesh) mf = a.mesh_f[(size_t)(f + t) * Sp + s];
            const f3 fc = force_at<COL>(x, fs, a, mf);
            U[t] = add(add(x, smul(a.dt, u)), smul(a.dt2, fc));
        };
        if (active) {
            if (!GATHER || gwin_ok) {
#pragma unroll
                for (uint32_t t = 0; t < kL; ++t) {
                    if (t >= cnt) continue;
                    f3 gv = mk(0.0f, 0.0f, 0.0f);
                    if (GATHER) {
                        const f3 q = grid_coord(X[t], origin, a.inv_h);
                        const Rng rg = point_range(q, G);
                        const bool inw = in_win(rg, gwin);
                        if (inw && rng_full(rg))
                            gv = gather_block(reinterpret_cast<const float4*>(win), gwin, rg, q);
                        else if (!rng_empty(rg))
                            gv = gather_any(reinterpret_cast<const float4*>(win), gwin, inw, a.gw.grid_f, rg, q, G);
                    }
                    predict(t, gv);
                }
            } else {
#pragma unroll
                for (uint32_t t = 0; t < kL; ++t) {
                    if (t >= cnt) continue;
                    const f3 q = grid_coord(X[t], origin, a.inv_h);
                    const Rng rg = point_range(q, G);
                    f3 gv = mk(0.0f, 0.0f, 0.0f);
                    if (rng_full(rg))
                        gv = gather_block(a.gw.grid_f, Win{0, 0, 0, G, G, G}, rg, q);  // the whole grid as the window
                    else if (!rng_empty(rg))
                        gv = gather_any_src<false>(nullptr, gwin, a.gw.grid_f, rg, q, G);
                    predict(t, gv);
                }
            }
        }
        if (GATHER && gwin_ok) __syncthreads();  // the window is reused below

        // ---------------- F: FTL, velocity, DFTL (the root-to-tip chain) ----------------
        Box nbox = box_empty();  // voxels of this segment's new positions (next step's gather)
        uint32_t RP[kL];         // packed voxel block of each new position (0: recompute)
#pragma unroll
        for (uint32_t t = 0; t < kL; ++t) RP[t] = 0;
        if (active) {
#pragma unroll
            for (uint32_t t = 0; t < kL; ++t) {
                if (t >= cnt) continue;
                const uint32_t bo = b0 + t * bS;
                const f3 x = X[t];
                const f3 pred = U[t];
                // follow the leader (compute.comp:191-192)
                const f3 pn = add(par, smul(r, normalize(sub(pred, par))));
                // damped velocity, clamp, correction (compute.comp:195-201);
                // dot > vmax2 is exactly sqrt(dot) > vmax
                f3 vel = smul(a.D, muls(sub(pn, x), a.inv_dt));
                if (dot(vel, vel) > a.vmax2) vel = muls(normalize(vel), a.vmax);
                const f3 corr = smul(a.D, sub(pn, pred));
                st3(posb, bo, P4, P8, pn);
                if (a.corr) st3(a.corr, bo, P4, P8, corr);
                // DFTL for point f+t-1 (compute.comp:213-215)
                const f3 dc = muls(corr, a.inv_dt);
                if (t == 0) {
                    if (f >= 2) {
                        cv = sub(cv, dc);
                        st3(velb, bo - bS, P4, P8, cv);
                    }
                } else {
                    U[t - 1] = sub(U[t - 1], dc);
                    st3(velb, bo - bS, P4, P8, U[t - 1]);
                }
                X[t] = pn;
                U[t] = vel;
                par = pn;
                if (SPLAT) {
                    const Rng rn = point_range(grid_coord(pn, origin, a.inv_h), G);
                    box_add(nbox, rn);
                    RP[t] = pack_block(rn);
                }
                if (t + 1 == cnt && last_seg) st3(velb, bo, P4, P8, vel);  // tip: no correction (compute.comp:213)
            }
        }

        // ---------------- S: splat of the finished points ----------------
        if (SPLAT) {
            // finished: point f-1 (carried), points f..f+cnt-2, and the tip
            Box sbox = box_empty();
            if (active) {
                if (f >= 2) box_add(sbox, point_range(grid_coord(cp, origin, a.inv_h), G));
#pragma unroll
                for (uint32_t t = 0; t < kL; ++t)
                    if (t + 1 < cnt || (t + 1 == cnt && last_seg))
                        box_add(sbox, RP[t] ? unpack_block(RP[t]) : point_range(grid_coord(X[t], origin, a.inv_h), G));
            }
            const Box gbn = block_box(nbox, red);
            if (threadIdx.x == 0) {
                int* gb = a.gw.gbox + ((size_t)wg * a.nseg + sg) * 6;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    gb[k] = gbn.lo[k];
                    gb[3 + k] = gbn.hi[k];
                }
            }
            const Box box = block_box(sbox, red);
            const Plan pl = plan_box(box, min((int)a.win_splat, (int)kWinCells), true);
            uint32_t* ctr = a.gw.counters + a.parity * 4;
            if (!box_is_empty(box) && (!pl.ok || pl.ny * pl.nz > (int)kMaxDeferredChunks)) {
                // no window fits (or more chunks than a list entry counts):
                // device atomics into the overflow grid
                const Win none{0, 0, 0, 0, 0, 0};
                const SplatLane sl = splat_lane((int)kWinCells);
                if (active) {
                    if (f >= 2) {
                        const f3 q = grid_coord(cp, origin, a.inv_h);
                        const Rng rg = point_range(q, G);
                        if (!rng_empty(rg)) splat_any(win, none, false, sl, a.gw.overflow, rg, q, cv, G);
                    }
#pragma unroll
                    for (uint32_t t = 0; t < kL; ++t) {
                        if (!(t + 1 < cnt || (t + 1 == cnt && last_seg))) continue;
                        const f3 q = grid_coord(X[t], origin, a.inv_h);
                        const Rng rg = point_range(q, G);
                        if (!rng_empty(rg)) splat_any(win, none, false, sl, a.gw.overflow, rg, q, U[t], G);
                    }
                }
                if (threadIdx.x == 0) atomicOr(ctr + 2, 1u);
            }
            // this workgroup-segment's own pool slot and descriptor
            const uint32_t ws = wg * a.nseg + sg;
            bool slot_used = false;
            if (pl.ok && pl.ny * pl.nz > 1 && pl.ny * pl.nz <= (int)kMaxDeferredChunks) {
                // A box larger than one window: splatting it here would take
                // one pass over the segment's points per chunk.  Publish the
                // plan instead; the segment's positions and final velocities
                // are already stored, so workgroups that have finished their
                // own strands can take chunks in parallel, and this workgroup
                // takes what is left once it is done (deferred_chunks).
                __threadfence();  // this thread's stored positions and velocities, before the entry
                __syncthreads();
                if (threadIdx.x == 0) own_k[n_own[0]++] = publish_deferred(a, ws, box, pl);
            } else if (pl.ok && pl.ny * pl.nz == 1) {
                // the whole box in one window
                const Win cw = chunk_of(box, pl, 0, 0);
                const int cells = cw.ex * cw.ey * cw.ez;
                const SplatLane sl = splat_lane(cells);
                for (int c = threadIdx.x; c < cells * sl.nc; c += kW) win[c] = make_int4(0, 0, 0, 0);
                __syncthreads();
                if (active) {
                    if (f >= 2) {
                        const f3 q = grid_coord(cp, origin, a.inv_h);
                        const Rng rg = point_range(q, G);
                        if (!rng_empty(rg)) {
                            if (rng_full(rg))
                                splat_block(win, cw, rg, q, cv, sl);
                            else
                                splat_any(win, cw, true, sl, a.gw.overflow, rg, q, cv, G);
                        }
                    }
#pragma unroll
                    for (uint32_t t = 0; t < kL; ++t) {
                        if (!(t + 1 < cnt || (t + 1 == cnt && last_seg))) continue;
                        const f3 q = grid_coord(X[t], origin, a.inv_h);
                        const Rng rg = RP[t] ? unpack_block(RP[t]) : point_range(q, G);
                        if (rng_empty(rg)) continue;
                        if (rng_full(rg))
                            splat_block(win, cw, rg, q, U[t], sl);
                        else
                            splat_any(win, cw, true, sl, a.gw.overflow, rg, q, U[t], G);
                    }
                }
                __syncthreads();
                dump_window(a, cw, cells, sl, (int)ws, win, bcast);
                slot_used = true;
            }
            if (!slot_used && threadIdx.x == 0) {
                // the slot's descriptor from two steps ago must not be read
                a.gw.desc[((size_t)a.parity * a.gw.desc_cap + ws) * kDescInts + 3] = 0;
            }
        }
        // carry the segment's last point into the next segment
#pragma unroll
        for (uint32_t t = 0; t < kL; ++t)
            if (t + 1 == cnt) {
                cp = X[t];
                cv = U[t];
            }
    }
    if (SPLAT) deferred_chunks(a, wg, win, bcast, reinterpret_cast<uint32_t*>(red), own_k, n_own);
}

template <bool GATHER, bool SPLAT, int COL>
__global__ __launch_bounds__(kW) void k_step(StepArgs a)
{
    const uint32_t wg = a.swz ? xcd_swizzle(blockIdx.x, a.nwg) : blockIdx.x;
    if (wg >= a.nwg) return;
    step_body<GATHER, SPLAT, COL>(a, wg);
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
    m.spheres = global_ptr(in.spheres);
    m.cell_start = global_ptr(in.cell_start);
    m.cell_tris = global_ptr(in.cell_tris);
    m.cell_spheres = global_ptr(in.cell_spheres);
    m.cell_keys = global_ptr(in.cell_keys);
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

template <bool GATHER, bool SPLAT, int COL>
__global__ __launch_bounds__(kW) void k_step_batch(const StepArgs* __restrict__ args,
                                                   const uint32_t* __restrict__ wg_start, uint32_t ngroom,
                                                   uint32_t swz)
{
    const uint32_t total = wg_start[ngroom];
    const uint32_t gwg = swz ? xcd_swizzle(blockIdx.x, total) : blockIdx.x;
    if (gwg >= total) return;
    const uint32_t g = find_groom(wg_start, ngroom, gwg);
    const StepArgs a = global_args(args[g]);
    step_body<GATHER, SPLAT, COL>(a, gwg - wg_start[g]);
}

// Mesh pass over many grooms: args[idx[m]] is the m-th groom with a mesh,
// blk_start[m] its first block (prefix sums, nm + 1 entries).  Arguments are
// read from device memory, so a captured launch follows later calls' values.
__global__ __launch_bounds__(256) void k_mesh_force_batch(const StepArgs* __restrict__ args,
                                                          const uint32_t* __restrict__ idx,
                                                          const uint32_t* __restrict__ blk_start, uint32_t nm)
{
    if (blockIdx.x >= blk_start[nm]) return;
    const uint32_t m = find_groom(blk_start, nm, blockIdx.x);
    const StepArgs& a = args[idx[m]];
    const uint32_t plane = a.N * a.Sp;
    const uint32_t i = (blockIdx.x - blk_start[m]) * 256u + threadIdx.x;
    if (i < plane)
        mesh_point(global_mesh(a.mesh), global_ptr(a.pos), plane, a.k, const_cast<float4*>(global_ptr(a.mesh_f)), i);
}

// ---------------------------------------------------------------------------
// Reduce: partial boxes -> grid (GridCell ints + float cells); clears the grid.
// ---------------------------------------------------------------------------
// One wavefront per brick of kBX x kBY x kBZ = 64 voxels, one voxel per lane:
// each lane sums its voxel over the boxes listed for the brick in registers
// (wrapping int32 adds, order independent), kInFlight boxes per round.
constexpr uint32_t kReduceThreads = 64 * kBricksPerWG;
constexpr uint32_t kInFlight = 16;
static_assert(kBX * kBY * kBZ == 64, "one voxel per lane");

struct BoxRef {
    int x0, y0, z0, ex, ey, ez, off;
};

__device__ __forceinline__ BoxRef box_ref(const int* __restrict__ d)
{
    return BoxRef{d[0], d[1], d[2], d[3], d[4], d[5], d[6]};
}

// This lane's cell of box b (a zero int4 when the voxel is outside it).
__device__ __forceinline__ int4 box_cell(const BoxRef& b, const int4* __restrict__ pool, int cx, int cy, int cz)
{
    const uint32_t ux = (uint32_t)(cx - b.x0), uy = (uint32_t)(cy - b.y0), uz = (uint32_t)(cz - b.z0);
    if (ux < (uint32_t)b.ex && uy < (uint32_t)b.ey && uz < (uint32_t)b.ez)
        return pool[(uint32_t)b.off + (uz * (uint32_t)b.ey + uy) * (uint32_t)b.ex + ux];
    return make_int4(0, 0, 0, 0);
}

__device__ __forceinline__ void add4(uint4& acc, int4 v)
{
    acc.x += (uint32_t)v.x;
    acc.y += (uint32_t)v.y;
    acc.z += (uint32_t)v.z;
    acc.w += (uint32_t)v.w;
}

__device__ __forceinline__ bool near_limit(int v) { return v >= (1 << 30) || v <= -(1 << 30); }

__device__ __forceinline__ void reduce_brick(const GridWork& gw, uint32_t G, uint32_t parity, uint32_t brick,
                                             bool reset_counters)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nbr = gw.nbx * gw.nby * gw.nbz;
    const uint32_t bx = brick % gw.nbx, by = (brick / gw.nbx) % gw.nby, bz = brick / (gw.nbx * gw.nby);
    const int cx = (int)(bx * kBX + (lane % kBX)), cy = (int)(by * kBY + (lane / kBX) % kBY),
              cz = (int)(bz * kBZ + lane / (kBX * kBY));
    const bool inside = cx < (int)G && cy < (int)G && cz < (int)G;
    const uint32_t* ctr = gw.counters + parity * 4;
    const bool ovf = ctr[2] != 0;
    const uint32_t nlisted = gw.brick_cnt[parity * nbr + brick];
    if (lane == 0) gw.brick_cnt[(parity ^ 1u) * nbr + brick] = 0;  // for the step after next
    if (reset_counters) {
        // the next step's counters, and the chunk claims of the previous
        // step's deferred list (the next step uses the same half)
        uint32_t* nxt = gw.counters + (parity ^ 1u) * 4;
        const uint32_t nd = nxt[3];
        for (uint32_t i = lane; i < nd && i < gw.nws; i += 64) gw.dclaim[(size_t)(parity ^ 1u) * gw.nws + i] = 0;
        if (lane == 0) {
            nxt[0] = 0;
            nxt[1] = 0;
            nxt[2] = 0;
            nxt[3] = 0;
        }
    }
    // nothing listed, no overflow, and zero already: the brick stays as it is
    if (nlisted == 0 && !ovf && gw.brick_nz[brick] == 0) return;

    const int* desc = gw.desc + (size_t)parity * gw.desc_cap * kDescInts;
    const int4* pool = gw.pool + (size_t)parity * gw.pool_cap;
    uint4 acc = make_uint4(0, 0, 0, 0);
    // a compile-time capacity: with a run-time one this kernel measured 1.8x slower
    if (nlisted <= kBrickList) {
        const uint32_t* bl = gw.brick_list + ((size_t)parity * nbr + brick) * kBrickList;
        for (uint32_t base = 0; base < nlisted; base += 64) {
            // lane l fetches box base+l's descriptor: one dependent load chain per 64 boxes
            const uint32_t m = min(64u, nlisted - base);
            int4 d0 = make_int4(0, 0, 0, 0), d1 = make_int4(0, 0, 0, 0);
            if (lane < m) {
                const int4* dp = reinterpret_cast<const int4*>(desc + (size_t)bl[base + lane] * kDescInts);
                d0 = dp[0];
                d1 = dp[1];
            }
            // then every lane walks the boxes, kInFlight cell loads at a time
            for (uint32_t j = 0; j < m; j += kInFlight) {
                int4 v[kInFlight];
#pragma unroll
                for (uint32_t u = 0; u < kInFlight; ++u) {
                    v[u] = make_int4(0, 0, 0, 0);
                    if (j + u < m) {
                        const int src = (int)(j + u);
                        const BoxRef b{__builtin_amdgcn_readlane(d0.x, src), __builtin_amdgcn_readlane(d0.y, src),
                                       __builtin_amdgcn_readlane(d0.z, src), __builtin_amdgcn_readlane(d0.w, src),
                                       __builtin_amdgcn_readlane(d1.x, src), __builtin_amdgcn_readlane(d1.y, src),
                                       __builtin_amdgcn_readlane(d1.z, src)};
                        v[u] = box_cell(b, pool, cx, cy, cz);
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < kInFlight; ++u) add4(acc, v[u]);
            }
        }
    } else {
        // the brick's list overflowed: scan every box of the step
        const uint32_t ndesc = min(gw.nws + ctr[1], gw.desc_cap);
        const int x0 = (int)(bx * kBX), y0 = (int)(by * kBY), z0 = (int)(bz * kBZ);
        for (uint32_t i = 0; i < ndesc; ++i) {
            const BoxRef b = box_ref(desc + (size_t)i * kDescInts);
            if (b.ex <= 0 || b.x0 > x0 + (int)kBX - 1 || b.x0 + b.ex <= x0 || b.y0 > y0 + (int)kBY - 1 ||
                b.y0 + b.ey <= y0 || b.z0 > z0 + (int)kBZ - 1 || b.z0 + b.ez <= z0)
                continue;
            add4(acc, box_cell(b, pool, cx, cy, cz));
        }
    }
    bool nz = false, near = false;
    if (inside) {
        const uint32_t gi = (uint32_t)cx + (uint32_t)cy * G + (uint32_t)cz * G * G;
        if (ovf) {
            add4(acc, gw.overflow[gi]);
            gw.overflow[gi] = make_int4(0, 0, 0, 0);
        }
        const int4 v = make_int4((int)acc.x, (int)acc.y, (int)acc.z, (int)acc.w);
        nz = (v.x | v.y | v.z | v.w) != 0;
        near = near_limit(v.x) || near_limit(v.y) || near_limit(v.z) || near_limit(v.w);
        gw.grid_i[gi] = v;
        gw.grid_f[gi] = make_float4((float)v.x, (float)v.y, (float)v.z, v.w > 0 ? 1.0f / (float)v.w : 0.0f);
    }
    const uint64_t nz_mask = __ballot(nz);
    const uint64_t near_mask = __ballot(near);
    if (lane == 0) {
        gw.brick_nz[brick] = nz_mask != 0 ? 1u : 0u;
        if (near_mask) atomicAdd(gw.near_limit, (unsigned long long)__popcll(near_mask));
    }
}

__global__ __launch_bounds__(kReduceThreads) void k_reduce(GridWork gw, uint32_t G, uint32_t parity)
{
    const uint32_t brick = blockIdx.x * kBricksPerWG + (threadIdx.x >> 6);
    if (brick >= gw.nbx * gw.nby * gw.nbz) return;
    reduce_brick(gw, G, parity, brick, brick == 0);
}

// Many grooms: brick_start[g] is the first global brick of groom g.
__global__ __launch_bounds__(kReduceThreads) void k_reduce_batch(const StepArgs* __restrict__ args,
                                                                 const uint32_t* __restrict__ brick_start,
                                                                 uint32_t ngroom)
{
    const uint32_t total = brick_start[ngroom];
    // one brick per wavefront: the index is wave-uniform, and saying so keeps
    // the groom lookup and its arguments in scalar registers
    const uint32_t gb = __builtin_amdgcn_readfirstlane(blockIdx.x * kBricksPerWG + (threadIdx.x >> 6));
    if (gb >= total) return;
    const uint32_t g = __builtin_amdgcn_readfirstlane(find_groom(brick_start, ngroom, gb));
    const uint32_t b = gb - brick_start[g];
    reduce_brick(global_grid(args[g].gw), args[g].G, args[g].parity, b, b == 0);
}

// ---------------------------------------------------------------------------
// AoS <-> SoA (reference Strand layout at the boundary)
// ---------------------------------------------------------------------------
// One thread per (physical strand slot, point).
__global__ void k_pack(const float4* __restrict__ aos, const uint32_t* __restrict__ perm, uint32_t S, uint32_t N,
                       uint32_t Sp, float* __restrict__ pos, float* __restrict__ vel, float* __restrict__ corr)
{
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (size_t)S * N) return;
    const uint32_t p = (uint32_t)(tid / N);
    const uint32_t i = (uint32_t)(tid % N);
    const uint32_t s = perm[p];
    const size_t plane = (size_t)N * Sp;
    const size_t base = (size_t)s * 3 * N;
    const float4 pv = aos[base + i];
    const float4 vv = aos[base + N + i];
    const size_t o = (size_t)i * Sp + p;
    pos[o] = pv.x;
    pos[plane + o] = pv.y;
    pos[2 * plane + o] = pv.z;
    vel[o] = vv.x;
    vel[plane + o] = vv.y;
    vel[2 * plane + o] = vv.z;
    if (corr) {
        const float4 c = aos[base + 2 * N + i];
        corr[o] = c.x;
        corr[plane + o] = c.y;
        corr[2 * plane + o] = c.z;
    }
}

// Output tiles: a workgroup stages ts consecutive strand slots (all N points)
// in LDS, reading each point plane with consecutive lanes on consecutive
// slots, then writes every strand's block contiguously in caller order.
// Rows are padded to N+1 entries against LDS bank conflicts.
constexpr uint32_t kTileF4 = 2048;  // float4 entries of the LDS tile (32 KiB)
constexpr uint32_t kTileThreads = 256;

// SoA -> AoS (Strand layout); when `g` (the float cells of the last step's
// grid) is given, applies the pending grid friction to velocities of points
// 1..N-1 (compute.comp:257-298) on the way.
__global__ __launch_bounds__(kTileThreads) void k_unpack(float4* __restrict__ aos, const uint32_t* __restrict__ perm,
                                                         uint32_t S, uint32_t N, uint32_t Sp,
                                                         const float* __restrict__ pos, const float* __restrict__ vel,
                                                         const float* __restrict__ corr, const float4* __restrict__ g,
                                                         uint32_t G, float ox, float oy, float oz, float inv_h,
                                                         float mu, float omu, uint32_t ts)
{
    __shared__ float4 tile[kTileF4];
    const uint32_t p0 = blockIdx.x * ts;
    const uint32_t ns = min(ts, S - p0);
    const uint32_t row = N + 1;
    const size_t plane = (size_t)N * Sp;
    for (uint32_t arr = 0; arr < 3; ++arr) {
        for (uint32_t e = threadIdx.x; e < ts * N; e += kTileThreads) {
            const uint32_t pl = e % ts, i = e / ts;
            if (pl >= ns) continue;
            const size_t o = (size_t)i * Sp + p0 + pl;
            float4 val = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (arr == 0) {
                val = make_float4(pos[o], pos[plane + o], pos[2 * plane + o], 1.0f);
            } else if (arr == 1) {
                f3 v = mk(vel[o], vel[plane + o], vel[2 * plane + o]);
                if (g && i >= 1) {
                    const f3 pp = mk(pos[o], pos[plane + o], pos[2 * plane + o]);
                    const f3 q = grid_coord(pp, mk(ox, oy, oz), inv_h);
                    const Rng rg = point_range(q, (int)G);
                    f3 gv = mk(0.0f, 0.0f, 0.0f);
                    const int Gi = (int)G;
                    if (rng_full(rg))
                        gv = gather_block(g, Win{0, 0, 0, Gi, Gi, Gi}, rg, q);  // the whole grid as the window
                    else if (!rng_empty(rg))
                        gv = gather_any_src<false>(nullptr, Win{0, 0, 0, 0, 0, 0}, g, rg, q, Gi);
                    v = friction(v, gv, mu, omu);
                }
                val = make_float4(v.x, v.y, v.z, 0.0f);
            } else if (corr) {
                val = make_float4(corr[o], corr[plane + o], corr[2 * plane + o], 0.0f);
            }
            tile[pl * row + i] = val;
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < ns * N; e += kTileThreads) {
            const uint32_t pl = e / N, i = e % N;
            aos[((size_t)perm[p0 + pl] * 3 + arr) * N + i] = tile[pl * row + i];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Render-facing output
// ---------------------------------------------------------------------------
// Curve of the hair tessellation shader (hair.tese:34-80, `func`, one strand):
// sample j of M at v = j/(M-1); segment floor(v*(N-1)) between points v1, v2
// with neighbours v0, v3 (mirrored at the ends: v1 + (v1 - v2), v2 + (v2 -
// v1)); Bezier control points b1 = v1 + (1/3)((v2 - v0)/2), b2 = v2 -
// (1/3)((v3 - v1)/2); de Casteljau with mix(a,b,t) = a*(1-t) + b*t.  At v = 1
// the shader would read one point past the strand; here it is the tip
// (last segment, t = 1).  Positions staged per tile as in k_unpack; samples
// written strand by strand.
__device__ __forceinline__ f3 mix3(f3 a, f3 b, float t)
{
    const float u = 1.0f - t;
    return mk(a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t);
}

__global__ __launch_bounds__(kTileThreads) void k_resample(const float* __restrict__ pos,
                                                           const uint32_t* __restrict__ perm, uint32_t S, uint32_t N,
                                                           uint32_t Sp, uint32_t M, float4* __restrict__ out,
                                                           uint32_t ts)
{
    __shared__ float4 tile[kTileF4];
    const uint32_t p0 = blockIdx.x * ts;
    const uint32_t ns = min(ts, S - p0);
    const uint32_t row = N + 1;
    const size_t plane = (size_t)N * Sp;
    for (uint32_t e = threadIdx.x; e < ts * N; e += kTileThreads) {
        const uint32_t pl = e % ts, i = e / ts;
        if (pl >= ns) continue;
        const size_t o = (size_t)i * Sp + p0 + pl;
        tile[pl * row + i] = make_float4(pos[o], pos[plane + o], pos[2 * plane + o], 0.0f);
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < ns * M; e += kTileThreads) {
        const uint32_t pl = e / M, j = e % M;
        const float4* T = tile + pl * row;
        auto P = [&](uint32_t i) {
            const float4 v = T[i];
            return mk(v.x, v.y, v.z);
        };
        const float v = (float)j / (float)(M - 1);
        const float x = v * (float)(N - 1);
        const float fl = floorf(x);
        uint32_t seg = (uint32_t)fl;
        float t = x - fl;
        if (seg >= N - 1) {
            seg = N - 2;
            t = 1.0f;
        }
        const f3 v1 = P(seg), v2 = P(seg + 1);
        const f3 v0 = seg == 0 ? add(v1, sub(v1, v2)) : P(seg - 1);
        const f3 v3 = seg + 2 == N ? add(v2, sub(v2, v1)) : P(seg + 2);
        const float third = 1.0f / 3.0f;
        const f3 b0 = v1;
        const f3 b1 = add(v1, smul(third, muls(sub(v2, v0), 0.5f)));
        const f3 b2 = sub(v2, smul(third, muls(sub(v3, v1), 0.5f)));
        const f3 b3 = v2;
        const f3 b01 = mix3(b0, b1, t), b11 = mix3(b1, b2, t), b21 = mix3(b2, b3, t);
        const f3 b02 = mix3(b01, b11, t), b12 = mix3(b11, b21, t);
        const f3 c = mix3(b02, b12, t);
        out[(size_t)perm[p0 + pl] * M + j] = make_float4(c.x, c.y, c.z, 1.0f);
    }
}

// Strand slots per output tile: whole padded rows of N + 1 entries.
static uint32_t tile_strands(uint32_t N) { return max(1u, min(64u, kTileF4 / (N + 1))); }

// The indirect draw record the shader's count produces (compute.comp:302).
__global__ void k_indirect(uint32_t* __restrict__ rec, uint32_t S)
{
    if (threadIdx.x < 4) rec[threadIdx.x] = threadIdx.x == 0 ? S : (threadIdx.x == 1 ? 1u : 0u);
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
template <bool G, bool S>
static void launch_step_e(const StepArgs& a, int col, uint32_t blocks, hipStream_t st)
{
    switch (col) {
    case 0: hipLaunchKernelGGL((k_step<G, S, 0>), dim3(blocks), dim3(kW), 0, st, a); break;
    case kColEll: hipLaunchKernelGGL((k_step<G, S, kColEll>), dim3(blocks), dim3(kW), 0, st, a); break;
    case kColMesh: hipLaunchKernelGGL((k_step<G, S, kColMesh>), dim3(blocks), dim3(kW), 0, st, a); break;
    default:
        hipLaunchKernelGGL((k_step<G, S, kColEll | kColMesh>), dim3(blocks), dim3(kW), 0, st, a);
        break;
    }
}

hipError_t launch_step(const StepArgs& a, bool gather, bool splat, hipStream_t st)
{
    const uint32_t blocks = ((a.nwg + 7) / 8) * 8;
    if (a.mesh.tris && !a.mesh_f) return hipErrorInvalidValue;
    const int col = (a.ncol > 1 ? kColEll : 0) | (a.mesh.tris ? kColMesh : 0);
    if (gather) {
        if (splat)
            launch_step_e<true, true>(a, col, blocks, st);
        else
            launch_step_e<true, false>(a, col, blocks, st);
    } else {
        if (splat)
            launch_step_e<false, true>(a, col, blocks, st);
        else
            launch_step_e<false, false>(a, col, blocks, st);
    }
    return hipGetLastError();
}

template <bool G, bool S>
static void launch_batch_e(const StepArgs* args, const uint32_t* wg_start, uint32_t ngroom, int col,
                           uint32_t swz, uint32_t blocks, hipStream_t st)
{
    switch (col) {
    case 0:
        hipLaunchKernelGGL((k_step_batch<G, S, 0>), dim3(blocks), dim3(kW), 0, st, args, wg_start, ngroom, swz);
        break;
    case kColEll:
        hipLaunchKernelGGL((k_step_batch<G, S, kColEll>), dim3(blocks), dim3(kW), 0, st, args, wg_start, ngroom, swz);
        break;
    case kColMesh:
        hipLaunchKernelGGL((k_step_batch<G, S, kColMesh>), dim3(blocks), dim3(kW), 0, st, args, wg_start, ngroom, swz);
        break;
    default:
        hipLaunchKernelGGL((k_step_batch<G, S, kColEll | kColMesh>), dim3(blocks), dim3(kW), 0, st, args, wg_start,
                           ngroom, swz);
        break;
    }
}

hipError_t launch_step_batch(const StepArgs* args_dev, const uint32_t* wg_start_dev, uint32_t ngroom,
                             uint32_t total_wg, bool gather, bool splat, int col, uint32_t swz, hipStream_t st)
{
    const uint32_t blocks = ((total_wg + 7) / 8) * 8;
    if (gather) {
        if (splat)
            launch_batch_e<true, true>(args_dev, wg_start_dev, ngroom, col, swz, blocks, st);
        else
            launch_batch_e<true, false>(args_dev, wg_start_dev, ngroom, col, swz, blocks, st);
    } else {
        if (splat)
            launch_batch_e<false, true>(args_dev, wg_start_dev, ngroom, col, swz, blocks, st);
        else
            launch_batch_e<false, false>(args_dev, wg_start_dev, ngroom, col, swz, blocks, st);
    }
    return hipGetLastError();
}

hipError_t launch_reduce_batch(const StepArgs* args_dev, const uint32_t* brick_start_dev, uint32_t ngroom,
                               uint32_t total_bricks, hipStream_t st)
{
    const uint32_t blocks = (total_bricks + kBricksPerWG - 1) / kBricksPerWG;
    hipLaunchKernelGGL(k_reduce_batch, dim3(blocks), dim3(kReduceThreads), 0, st, args_dev, brick_start_dev, ngroom);
    return hipGetLastError();
}

hipError_t launch_mesh_force(const StepArgs& a, hipStream_t st)
{
    const uint32_t plane = a.N * a.Sp;
    if (!a.mesh.tris || !a.mesh_f) return hipErrorInvalidValue;
    float4* out = const_cast<float4*>(a.mesh_f);
    hipLaunchKernelGGL(k_mesh_force, dim3((plane + 255) / 256), dim3(256), 0, st, a.mesh, a.pos, plane, a.k, out);
    return hipGetLastError();
}

hipError_t launch_mesh_force_batch(const StepArgs* args_dev, const uint32_t* idx_dev, const uint32_t* blk_start_dev,
                                   uint32_t nm, uint32_t total_blocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_mesh_force_batch, dim3(total_blocks), dim3(256), 0, st, args_dev, idx_dev, blk_start_dev, nm);
    return hipGetLastError();
}

hipError_t launch_reduce(const GridWork& gw, uint32_t G, uint32_t parity, hipStream_t st)
{
    (void)G;
    const uint32_t nbr = gw.nbx * gw.nby * gw.nbz;
    hipLaunchKernelGGL(k_reduce, dim3((nbr + kBricksPerWG - 1) / kBricksPerWG), dim3(kReduceThreads), 0, st, gw, G,
                       parity);
    return hipGetLastError();
}

hipError_t launch_pack(const float* aos_dev, const uint32_t* perm, uint32_t S, uint32_t N, uint32_t Sp, float* pos,
                       float* vel, float* corr, hipStream_t st)
{
    const size_t n = (size_t)S * N;
    const uint32_t bs = 256;
    const size_t blocks = (n + bs - 1) / bs;
    hipLaunchKernelGGL(k_pack, dim3((uint32_t)blocks), dim3(bs), 0, st, reinterpret_cast<const float4*>(aos_dev),
                       perm, S, N, Sp, pos, vel, corr);
    return hipGetLastError();
}

hipError_t launch_resample(const float* pos, const uint32_t* perm, uint32_t S, uint32_t N, uint32_t Sp, uint32_t M,
                           float* out_dev, hipStream_t st)
{
    if (N < 2 || M < 2 || N + 1 > kTileF4) return hipErrorInvalidValue;
    const uint32_t ts = tile_strands(N);
    hipLaunchKernelGGL(k_resample, dim3((S + ts - 1) / ts), dim3(kTileThreads), 0, st, pos, perm, S, N, Sp, M,
                       reinterpret_cast<float4*>(out_dev), ts);
    return hipGetLastError();
}

hipError_t launch_indirect(uint32_t* rec_dev, uint32_t S, hipStream_t st)
{
    hipLaunchKernelGGL(k_indirect, dim3(1), dim3(64), 0, st, rec_dev, S);
    return hipGetLastError();
}

hipError_t launch_unpack(float* aos_dev, const uint32_t* perm, uint32_t S, uint32_t N, uint32_t Sp,
                         const float* pos, const float* vel, const float* corr, const UnpackGrid* g, hipStream_t st)
{
    if (N + 1 > kTileF4) return hipErrorInvalidValue;
    const uint32_t ts = tile_strands(N);
    const dim3 grid((S + ts - 1) / ts), block(kTileThreads);
    if (g)
        hipLaunchKernelGGL(k_unpack, grid, block, 0, st, reinterpret_cast<float4*>(aos_dev), perm, S, N, Sp, pos, vel,
                           corr, g->grid, g->G, g->ox, g->oy, g->oz, g->inv_h, g->mu, g->omu, ts);
    else
        hipLaunchKernelGGL(k_unpack, grid, block, 0, st, reinterpret_cast<float4*>(aos_dev), perm, S, N, Sp, pos, vel,
                           corr, (const float4*)nullptr, 0u, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f, ts);
    return hipGetLastError();
}

}  // namespace hs
