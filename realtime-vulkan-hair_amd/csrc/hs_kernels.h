// hs_kernels.h - launch interface between the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hs_device_math.hpp"

namespace hs {

// Tile shapes the step kernels are built for (hs_step.hip, once per shape):
// strands per workgroup (one lane each) x points per segment of the
// root-to-tip sweep.  Results are identical for every shape.  Strands are
// ordered in k-d leaves of kLeafStrands; a workgroup covers whole leaves, and
// the strand planes are padded to a multiple of kMaxTileStrands so every shape
// fits the same layout.
struct TileShape {
    uint32_t strands, points;
};
constexpr TileShape kShapes[] = {{128, 16}, {256, 8}, {256, 12}};
constexpr int kNumShapes = 3;
constexpr uint32_t kLeafStrands = 128;
constexpr uint32_t kMaxTileStrands = 256;
#ifndef HS_WINDOW_CELLS
#define HS_WINDOW_CELLS 2048
#endif
constexpr uint32_t kWinCells = HS_WINDOW_CELLS;  // LDS window of one workgroup, 16 B entries
// reduce kernel: bricks of kBX x kBY x kBZ voxels, one wavefront each
constexpr uint32_t kBX = 8, kBY = 4, kBZ = 2;
constexpr uint32_t kBricksPerWG = 4;
constexpr uint32_t kMaxColliders = 8;
constexpr uint32_t kDescInts = 8;     // partial-box descriptor: lo xyz, ext xyz, pool offset, pad
constexpr uint32_t kBrickList = 2048;  // box indices kept per reduce brick (more: it scans every box)
constexpr int kDBoxInts = 12;          // deferred segment: box lo xyz, hi xyz, chunk extents y z, counts y z, pad
constexpr uint32_t kSlotCells = HS_WINDOW_CELLS;  // pool cells a workgroup-segment owns for its
                                                 // first box (a whole splat window)

// Device-side grid workspace of one groom (all zeroed at creation).
struct GridWork {
    int4* grid_i;         // G^3 GridCell ints: the reference-visible grid of the last step
    float4* grid_f;       // G^3 (float(vx), float(vy), float(vz), density > 0 ? 1/float(density) : 0)
    int4* pool;           // partial boxes written by the step kernel, [2][pool_cap]
    int* desc;            // partial-box descriptors, [2][desc_cap][kDescInts]
    uint32_t* counters;   // [2][4]: shared cells used, shared boxes used, fallback flag, deferred entries
    int4* overflow;       // G^3 int4 atomic fallback when the pool is full (normally zero)
    int* gbox;            // [nwg][nseg][6] voxel box each workgroup-segment touched last step (laid
                          // out for the last step's tile shape)
    uint32_t* brick_cnt;  // [2][nbricks] boxes listed per brick (may exceed kBrickList)
    uint32_t* brick_list; // [2][nbricks][kBrickList] indices of the boxes overlapping a brick
    uint32_t* brick_nz;   // [nbricks] 1 when the brick's grid cells are not all zero
    unsigned long long* near_limit;  // (cell, step) pairs with a component >= 2^30 in magnitude
    // workgroup-segments whose box needs several windows, splatted chunk by
    // chunk by any workgroup of the groom that is done with its own strands
    unsigned long long* dlist;  // [2][nws_cap] step tag (28 bits) | chunks (12) | workgroup-segment (24)
    uint32_t* dclaim;           // [2][nws_cap] chunks taken of each list entry
    int* dbox;                  // [nws_cap][kDBoxInts] box and chunking of a deferred segment
    // pool / descriptors per parity: first each workgroup-segment's own slot
    // (kSlotCells cells, descriptor index = its number), then a shared part
    // taken with atomics; pool_cap / desc_cap count both
    uint32_t pool_cap, desc_cap;
    uint32_t nws;            // workgroup-segments of the launch's tile shape: owners of the slots
    uint32_t nws_cap;        // the most any tile shape has: the stride of dlist / dclaim halves
    uint32_t nbx, nby, nbz;  // reduce bricks per axis
};

// Everything one step launch needs, passed by value as the kernel argument.
// Triangle-mesh collider: per-triangle records of 10 float4 (a, b, c, face
// normal, edge pseudo-normals ab bc ca, vertex pseudo-normals a b c) and a
// uniform grid of candidate lists.
struct MeshArgs {
    const float4* tris;
    const uint32_t* cell_start;  // [cells + 1]
    // [entries][3]: (a, key), (b, triangle id bits), (c, bounding radius);
    // key = a lower bound of the distance from the cell's centre to the
    // triangle, ascending within a cell; the bounding centre is
    // ((a + b) + c) * (1/3) in binary32
    const float4* cell_recs;
    float half_diag;             // half the cell diagonal
    float ox, oy, oz;            // grid origin
    float icx, icy, icz;         // 1 / cell size
    uint32_t dx, dy, dz;         // cells per axis
    float margin2;               // contact when |x - q|^2 < margin2
};

// Collider mode of a step kernel: which colliders beyond the sphere it handles.
constexpr int kColEll = 1;
constexpr int kColMesh = 2;  // mesh contact read from the mesh pass's result (StepArgs::mesh_f)

struct StepArgs {
    float* pos;           // [3][N][Sp]
    float* vel;           // [3][N][Sp]
    float* corr;          // [3][N][Sp] or null (correction vectors not kept)
    const float* seg;     // [Sp] segment rest length
    const float* noise;   // [Sp] per-strand wind noise or null
    uint32_t S, N, Sp, nwg, nseg;
    // voxel grid (compute.comp:9-11, 204-206)
    GridWork gw;
    uint32_t G, parity;   // parity selects pool/counter half written by this launch
    uint32_t tag;         // 1 + steps of this groom launched before, mod 2^28 - 1 (tells deferred-list
                          // entries apart; never 0, the value of a list slot not written yet)
    float ox, oy, oz, inv_h;
    float mu_prev, omu_prev;  // friction of the previous step and 1 - friction
    // step constants
    float dt, inv_dt, dt2, D, k, vmax, sphere_r;
    float vmax2;      // largest x with sqrt(x) <= vmax
    float sphere_r2;  // smallest x with sqrt(x) >= sphere_r
    float fx, fy, fz;     // gravity + uniform wind
    float ndx, ndy, ndz;  // per-strand noise direction
    float sphere_c[3];    // collider 0 centre (Collider.transform[3].xyz)
    uint32_t ncol;        // colliders in use (0: none, 1: the sphere, >1: + ellipsoids)
    const Collider* cols; // device copy of all ncol colliders (read for ellipsoids 1..ncol-1)
    MeshArgs mesh;        // triangle-mesh collider (mesh.tris null: none)
    const float4* mesh_f; // [N][Sp] mesh contact of the stored positions (launch_mesh_force), or null
    uint32_t forcing;     // wind forcing mode (0: none); its per-step constants:
    float f_amp, f_a2, f_t3;
    // tile choices (results do not depend on them)
    uint32_t win_gather;  // voxel cells of the gather window (<= kWinCells)
    uint32_t win_splat;   // voxel cells of a splat window (<= kWinCells)
    uint32_t swz;         // 1: XCD-aware workgroup order
    uint32_t gbox_ok;     // 1: gw.gbox holds the last step's boxes in this launch's tile shape
};

struct UnpackGrid {
    const float4* grid;  // float cells (GridWork::grid_f) of the grid to gather from
    uint32_t G;
    float ox, oy, oz, inv_h, mu, omu;
};

// Step kernels of each tile shape (namespace t<strands>x<points>).
#define HS_SHAPE_LAUNCHERS(NS)                                                                                   \
    namespace NS {                                                                                               \
    hipError_t launch_step(const StepArgs& a, bool gather, bool splat, hipStream_t st);                          \
    /* many grooms per launch: per-groom arguments in device memory; wg_start is the prefix sum (ngroom + 1) */ \
    hipError_t launch_step_batch(const StepArgs* args_dev, const uint32_t* wg_start_dev, uint32_t ngroom,        \
                                 uint32_t total_wg, bool gather, bool splat, int col, uint32_t swz,              \
                                 hipStream_t st);                                                                \
    }
HS_SHAPE_LAUNCHERS(t128x16)
HS_SHAPE_LAUNCHERS(t256x8)
HS_SHAPE_LAUNCHERS(t256x12)
#undef HS_SHAPE_LAUNCHERS

// The step launchers of kShapes[shape].
inline hipError_t launch_step(int shape, const StepArgs& a, bool gather, bool splat, hipStream_t st)
{
    switch (shape) {
    case 1: return t256x8::launch_step(a, gather, splat, st);
    case 2: return t256x12::launch_step(a, gather, splat, st);
    default: return t128x16::launch_step(a, gather, splat, st);
    }
}
inline hipError_t launch_step_batch(int shape, const StepArgs* args_dev, const uint32_t* wg_start_dev,
                                    uint32_t ngroom, uint32_t total_wg, bool gather, bool splat, int col,
                                    uint32_t swz, hipStream_t st)
{
    switch (shape) {
    case 1: return t256x8::launch_step_batch(args_dev, wg_start_dev, ngroom, total_wg, gather, splat, col, swz, st);
    case 2: return t256x12::launch_step_batch(args_dev, wg_start_dev, ngroom, total_wg, gather, splat, col, swz, st);
    default: return t128x16::launch_step_batch(args_dev, wg_start_dev, ngroom, total_wg, gather, splat, col, swz, st);
    }
}
hipError_t launch_reduce(const GridWork& gw, uint32_t G, uint32_t parity, hipStream_t st);
// Mesh pass: mesh contact of every stored position of `a` into a.mesh_f
// (N * Sp float4).  Runs before a step kernel that reads it.
hipError_t launch_mesh_force(const StepArgs& a, hipStream_t st);
// The same for many grooms: args_dev[idx_dev[m]] for m < nm; blk_start_dev
// holds prefix sums of ceil(N * Sp / 256) blocks (nm + 1 entries).
hipError_t launch_mesh_force_batch(const StepArgs* args_dev, const uint32_t* idx_dev, const uint32_t* blk_start_dev,
                                   uint32_t nm, uint32_t total_blocks, hipStream_t st);
// Reduce of many grooms per launch: brick_start is the prefix sum (ngroom + 1
// entries) of their bricks.
hipError_t launch_reduce_batch(const StepArgs* args_dev, const uint32_t* brick_start_dev, uint32_t ngroom,
                               uint32_t total_bricks, hipStream_t st);
hipError_t launch_pack(const float* aos_dev, const uint32_t* perm, uint32_t S, uint32_t N, uint32_t Sp,
                       float* pos, float* vel, float* corr, hipStream_t st);
// Tessellation-shader curve samples (M per strand, float4 w = 1) in caller
// strand order; the indirect draw record {S, 1, 0, 0}.
hipError_t launch_resample(const float* pos, const uint32_t* perm, uint32_t S, uint32_t N, uint32_t Sp, uint32_t M,
                           float* out_dev, hipStream_t st);
hipError_t launch_indirect(uint32_t* rec_dev, uint32_t S, hipStream_t st);
hipError_t launch_unpack(float* aos_dev, const uint32_t* perm, uint32_t S, uint32_t N, uint32_t Sp,
                         const float* pos, const float* vel, const float* corr, const UnpackGrid* g,
                         hipStream_t st);

}  // namespace hs
