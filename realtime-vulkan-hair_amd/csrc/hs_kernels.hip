// hs_kernels.hip - HIP kernels of the strand-physics pass for gfx950.
//
// State layout in HBM (strand-major SoA, one plane per component):
//   pos[c][i][p], vel[c][i][p], c in {x,y,z}, i = point 0..N-1, p = physical
//   strand slot 0..Sp-1.  Strands are stored in k-d leaf order of their roots
//   (perm[p] = caller's strand index), so the strands of one workgroup are
//   spatial neighbours.  At a fixed point index a wavefront reads 64
//   consecutive floats of each plane: one 256 B line per plane.
//
// One step = two kernels.
//  k_step   (one workgroup per kW strands, hs_step.hip) walks each strand root
//           to tip in segments of kL points held in registers.  Per segment:
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
//
// This file holds the shape-independent kernels (mesh pass, reduce, layout
// conversion, resampling); the step kernels are in hs_step.hip, built once
// per tile shape.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device.hpp"

namespace hs {

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

// Marks a loaded value as used here (an empty asm statement that reads and
// may rewrite its registers; it emits no instruction).  Placed after a round's
// loads, it keeps the compiler from sinking a load into the branch that uses
// it, so all of them are issued before the first wait.
__device__ __forceinline__ void loaded_here(float4& v)
{
    asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
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
constexpr uint32_t kMeshRound = 2;  // mesh candidates loaded per round

__device__ __forceinline__ float4 mesh_contact(const MeshArgs& M, float px, float py, float pz, float k)
{
    const f3 x = mk(px, py, pz);
    const float4 none = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float gx = (x.x - M.ox) * M.icx, gy = (x.y - M.oy) * M.icy, gz = (x.z - M.oz) * M.icz;
    // outside the grid (or NaN): farther than the margin from every triangle.
    // Non-short-circuit '&', so the three coordinate loads are issued together.
    const bool inside = (gx >= 0.0f) & (gx < (float)M.dx) & (gy >= 0.0f) & (gy < (float)M.dy) & (gz >= 0.0f) &
                        (gz < (float)M.dz);
    if (!inside) return none;
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
    // candidates kMeshRound at a time: a candidate is one 48 B record holding
    // its key, bounding radius, id and vertices, so every load of a round is
    // in flight together and none waits on an id.  The bounding centre is
    // formed from the vertices, as the host formed it for the radius.  The
    // list is sorted by `key` (the distance from the cell's centre to the
    // triangle's bounding sphere, rounded down); key - |x - centre| bounds the
    // distance from below, so once that passes `reach` no later entry can
    // win.  The nearest of equal distances is the lowest id (the order of the
    // list does not matter).
    bool more = true;
    for (uint32_t i = i0; i < i1 && more; i += kMeshRound) {
        float4 ra[kMeshRound], rb[kMeshRound], rc[kMeshRound];
#pragma unroll
        for (uint32_t u = 0; u < kMeshRound; ++u) {
            const float4* r = M.cell_recs + 3 * (size_t)min(i + u, i1 - 1);
            ra[u] = r[0];
            rb[u] = r[1];
            rc[u] = r[2];
        }
#pragma unroll
        for (uint32_t u = 0; u < kMeshRound; ++u) {
            loaded_here(ra[u]);
            loaded_here(rb[u]);
            loaded_here(rc[u]);
        }
#pragma unroll
        for (uint32_t u = 0; u < kMeshRound; ++u) {
            if (i + u >= i1 || !more) break;
            if (ra[u].w - dxc > reach) {
                more = false;
                break;
            }
            const float third = 1.0f / 3.0f;
            const f3 cen = mk(((ra[u].x + rb[u].x) + rc[u].x) * third, ((ra[u].y + rb[u].y) + rc[u].y) * third,
                              ((ra[u].z + rb[u].z) + rc[u].z) * third);
            const f3 dc = sub(x, cen);
            const float lim = rc[u].w + reach;
            if (dot(dc, dc) > lim * lim * 1.0001f + 1e-12f) continue;
            const uint32_t t = __float_as_uint(rb[u].w);
            int feat;
            const f3 q = closest_on_tri(x, mk(ra[u].x, ra[u].y, ra[u].z), mk(rb[u].x, rb[u].y, rb[u].z),
                                        mk(rc[u].x, rc[u].y, rc[u].z), &feat);
            const f3 dq = sub(x, q);
            const float d2 = dot(dq, dq);
            if (d2 < best || (d2 == best && best_t != 0xffffffffu && t < best_t)) {
                best = d2;
                reach = sqrtf(best) + slack;
                best_t = t;
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
// kernel, whose register budget holds it to two or three waves per SIMD, this
// runs at full occupancy (8), which hides the scattered record loads; what
// bounds it is the bytes its lanes load, returned at a fixed rate per CU.
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
        for (uint32_t i = lane; i < nd && i < gw.nws_cap; i += 64) gw.dclaim[(size_t)(parity ^ 1u) * gw.nws_cap + i] = 0;
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
    // ts (N + 1) <= kTileF4: indices and divisors stay below 2^16, where the
    // reciprocal multiply divides exactly
    const uint32_t m_ts = div_magic(ts), m_n = div_magic(N);
    for (uint32_t arr = 0; arr < 3; ++arr) {
        for (uint32_t e = threadIdx.x; e < ts * N; e += kTileThreads) {
            const uint32_t i = div_small(e, ts, m_ts), pl = e - i * ts;
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
            const uint32_t pl = div_small(e, N, m_n), i = e - pl * N;
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
