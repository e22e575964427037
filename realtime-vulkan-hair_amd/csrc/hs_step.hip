// hs_step.hip - the step kernels (k_step, k_step_batch) for one tile shape.
//
// A tile shape is the workgroup's strands (kW, one lane each) and the points
// of one segment of the root-to-tip walk (kL, held in registers).  This file
// is compiled once per shape the library offers, each build in its own
// namespace (HS_SHAPE_NS) with HS_TILE_STRANDS / HS_TILE_POINTS set; results
// are the same bits for every shape, only the speed differs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device.hpp"

#ifndef HS_TILE_STRANDS
#define HS_TILE_STRANDS 128
#endif
#ifndef HS_TILE_POINTS
#define HS_TILE_POINTS 16
#endif
#ifndef HS_SHAPE_NS
#define HS_SHAPE_NS t128x16
#endif

namespace hs {
namespace HS_SHAPE_NS {

constexpr uint32_t kW = HS_TILE_STRANDS;  // strands per step workgroup
constexpr uint32_t kL = HS_TILE_POINTS;   // points per segment of the root-to-tip sweep

static_assert((uint64_t)kW * (kL + 1) * 1000000ull < (1ull << 32),
              "packed (vx<<32 | density) LDS sums need the density of one window below 2^32");
static_assert(kW % 64 == 0, "workgroup is whole wavefronts");
static_assert(kW % kLeafStrands == 0, "a workgroup is whole k-d leaves");
static_assert(kW <= kMaxTileStrands, "the strand planes are padded to kMaxTileStrands");
static_assert(kL <= 31, "per-point flags live in one 32-bit mask");

// segments of the longest strand (256 points): the most deferred-list
// entries one workgroup can publish in a step
#define HS_MAX_OWN_DEFERRED ((256 - 1 + HS_TILE_POINTS - 1) / HS_TILE_POINTS)

// ---------------------------------------------------------------------------
// Workgroup reductions and window planning
// ---------------------------------------------------------------------------
constexpr int kWaves = kW / 64;

// Minimum over the wavefront (every lane active): DPP steps within each row
// of 16 lanes (VALU moves, no LDS traffic), then the four row results.
__device__ __forceinline__ int wave_min(int v)
{
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

// Reduce two boxes over the workgroup in one pass (one pair of barriers);
// every thread gets both results.  `scratch` holds kWaves * 12 ints.
__device__ __forceinline__ void block_box2(const Box& b0, const Box& b1, int* scratch, Box* r0, Box* r1)
{
    int v[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = b0.lo[k];
        v[k + 3] = b0.hi[k] == INT32_MIN ? INT32_MAX : -b0.hi[k];
        v[k + 6] = b1.lo[k];
        v[k + 9] = b1.hi[k] == INT32_MIN ? INT32_MAX : -b1.hi[k];
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = wave_min(v[k]);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 12; ++k) scratch[wave * 12 + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        int m = scratch[k];
        for (int w = 1; w < kWaves; ++w) m = min(m, scratch[w * 12 + k]);
        v[k] = m;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r0->lo[k] = v[k];
        r0->hi[k] = v[k + 3] == INT32_MAX ? INT32_MIN : -v[k + 3];
        r1->lo[k] = v[k + 6];
        r1->hi[k] = v[k + 9] == INT32_MAX ? INT32_MIN : -v[k + 9];
    }
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

__device__ __forceinline__ Plan plan_box(const Box& b, int cap)
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
    __hip_atomic_store(a.gw.dlist + (size_t)a.parity * a.gw.nws_cap + k, dentry(a.tag, (uint32_t)(pl.ny * pl.nz), ws),
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
    const unsigned long long* list = a.gw.dlist + (size_t)a.parity * a.gw.nws_cap;
    uint32_t* claim = a.gw.dclaim + (size_t)a.parity * a.gw.nws_cap;
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
    const unsigned long long* list = a.gw.dlist + (size_t)a.parity * a.gw.nws_cap;
    const uint32_t* claim = a.gw.dclaim + (size_t)a.parity * a.gw.nws_cap;
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
    __shared__ int red[kW > kWaves * 12 ? kW : kWaves * 12];  // box reductions; the deferred-list scan
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
            if (active & (t < cnt)) {
                X[t] = ld3(posb, b0 + t * bS, P4, P8);
                U[t] = ld3(velb, b0 + t * bS, P4, P8);
            }
        }

        // ---------------- G: stage the previous step's voxel window ----------------
        Win gwin{0, 0, 0, 0, 0, 0};
        bool gwin_ok = false;
        // (after a change of tile shape the boxes are laid out for the old
        // shape: the points then gather straight from the float grid)
        if (GATHER && a.gbox_ok) {
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
                const uint32_t ex = (uint32_t)gwin.ex, ey = (uint32_t)gwin.ey;
                const uint32_t mx = div_magic(ex), my = div_magic(ey);
                for (int c0 = 0; c0 < cells; c0 += 8 * kW) {
                    // unconditional (clamped) loads: all 8 in flight before the stores
                    float4 tmp[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const uint32_t c = (uint32_t)min(c0 + k * (int)kW + (int)threadIdx.x, cells - 1);
                        const uint32_t rest = div_small(c, ex, mx);
                        const uint32_t ax = c - rest * ex;
                        const uint32_t az = div_small(rest, ey, my);
                        const uint32_t ay = rest - az * ey;
                        const uint32_t gi = (uint32_t)gwin.x0 + ax + ((uint32_t)gwin.y0 + ay) * (uint32_t)G +
                                            ((uint32_t)gwin.z0 + az) * GG;
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
                        if (inw & rng_full(rg))
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
            Box gbn, box;
            block_box2(nbox, sbox, red, &gbn, &box);
            if (threadIdx.x == 0) {
                int* gb = a.gw.gbox + ((size_t)wg * a.nseg + sg) * 6;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    gb[k] = gbn.lo[k];
                    gb[3 + k] = gbn.hi[k];
                }
            }
            const Plan pl = plan_box(box, min((int)a.win_splat, (int)kWinCells));
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

// Wavefronts per SIMD the compiler budgets registers for: 4 for the 256 x 8
// kernels with the mesh collider (they run on large meshes of many
// workgroups, where four wavefronts outweigh the few registers spilled), the
// occupancy the registers allow otherwise.
template <int COL>
constexpr int kStepWaves = (COL & kColMesh) && kW == 256 && kL == 8 ? 4 : 1;

template <bool GATHER, bool SPLAT, int COL>
__global__ __launch_bounds__(kW) __attribute__((amdgpu_waves_per_eu(kStepWaves<COL>))) void k_step(StepArgs a)
{
    const uint32_t wg = a.swz ? xcd_swizzle(blockIdx.x, a.nwg) : blockIdx.x;
    if (wg >= a.nwg) return;
    step_body<GATHER, SPLAT, COL>(a, wg);
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

}  // namespace HS_SHAPE_NS
}  // namespace hs
