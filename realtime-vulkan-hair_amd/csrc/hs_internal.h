// hs_internal.h - library-internal state shared by the C ABI (hs_api.cpp)
// and the job server (hs_server.cpp): groom layout, device state, context.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/hairsim.h"
#include "hs_kernels.h"

namespace hsi {

inline uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }
inline uint64_t grid_cells(uint32_t G) { return (uint64_t)G * G * G; }
// bricks of the reduce kernel along x, y, z
inline uint32_t bricks_x(uint32_t G) { return (G + hs::kBX - 1) / hs::kBX; }
inline uint32_t bricks_y(uint32_t G) { return (G + hs::kBY - 1) / hs::kBY; }
inline uint32_t bricks_z(uint32_t G) { return (G + hs::kBZ - 1) / hs::kBZ; }
inline uint64_t bricks(uint32_t G) { return (uint64_t)bricks_x(G) * bricks_y(G) * bricks_z(G); }

// Where a groom's device block comes from: hipMalloc when no heap is given,
// else a caller-owned heap (the job server's HBM arena).  release() is only
// called once the GPU no longer uses the block.
struct DevHeap {
    virtual ~DevHeap() = default;
    virtual hipError_t alloc(void** p, size_t bytes) = 0;
    virtual void release(void* p) = 0;
};

// Shape-derived sizes of one groom's device state.
// Strand planes padded to Sp slots; the grid workspace holds the pool slots,
// descriptors and deferred-list entries of the tile shape with the most
// workgroup-segments (nws_cap).
struct Layout {
    uint32_t Sp, nws_cap;
    uint64_t pool_cap, desc_cap;
};

// Workgroups and segments of an S x N groom in tile shape hs::kShapes[shape].
struct ShapeDims {
    uint32_t nwg, nseg;
};
ShapeDims shape_dims(int shape, uint32_t S, uint32_t N);

Layout layout_of(uint32_t S, uint32_t N, uint32_t G);

struct Groom {
    uint32_t S = 0, N = 0, G = 0, flags = 0;
    bool has_noise = false;
    Layout L{};
    float origin[3] = {0, 0, 0};
    float h = 1.0f;
    float noise_dir[3] = {0, 0, 0};
    float* pos = nullptr;
    float* vel = nullptr;
    float* corr = nullptr;
    float* seg = nullptr;
    float* noise = nullptr;
    uint32_t* perm = nullptr;      // physical slot -> caller's strand index
    hs::Collider* cols = nullptr;  // device copy of the step's colliders (ellipsoids)
    hs::GridWork gw{};
    void* block = nullptr;         // the one device block every array above lives in
    DevHeap* heap = nullptr;       // its source (null: hipMalloc)
    // Grid friction of the last step, not yet applied to the stored velocities.
    bool pending = false;
    float pend_mu = 0.0f, pend_omu = 1.0f;
    uint32_t parity = 0;           // pool/counter half the next splat writes
    uint32_t steps = 0;            // steps launched (the kernels' list tag)
    int last_shape = -1;           // tile shape of the last splat step (its gbox layout)
    uint64_t bytes = 0;            // size of `block`
};

// Size of a groom's device block; with `base`, also points g's arrays into it.
uint64_t carve(Groom& g, uint8_t* base);

// Checks a groom description; on failure returns the status and sets *why.
int check_groom_desc(const hs_groom_desc* d, const float* aos, std::string* why);

// Builds a groom's device state from a reference-layout buffer on stream
// `st` (block from `heap`, or hipMalloc), and waits for that stream.
int groom_build(const hs_groom_desc* d, const float* aos, hipStream_t st, DevHeap* heap, Groom* out,
                std::string* why);

// Returns the groom's block to its source.  The GPU must be done with it.
void groom_release(Groom& g);

// Reference-layout copy of the groom (pending friction applied) into device
// memory `aos_dev` (48*S*N bytes), on stream `st`.
hipError_t groom_unpack(const Groom& g, float* aos_dev, hipStream_t st);

// A mesh collider on the device (hs_mesh_create).
struct Mesh {
    void* block = nullptr;
    uint64_t bytes = 0;
    uint32_t ntris = 0;
    hs::MeshArgs args{};
};
}  // namespace hsi

struct hs_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<uint32_t, hsi::Groom> grooms;
    uint32_t next_id = 1;
    uint64_t bytes = 0;
    std::map<uint32_t, hsi::Mesh> meshes;
    uint32_t next_mesh = 1;
    // kernel timing: pooled events (step start, step end, reduce end, mesh
    // pass start) per step
    bool profile = false;
    struct Ev {
        hipEvent_t e[4];
        bool reduce, mesh;
    };
    std::vector<Ev> events;
    size_t events_used = 0;
    uint64_t prof_launches = 0;
    double prof_step_ms = 0.0, prof_reduce_ms = 0.0, prof_mesh_ms = 0.0;
    // hs_step with a mesh: the mesh pass's per-point result (grow-only)
    void* mesh_f = nullptr;
    size_t mesh_f_cap = 0;
    // hs_step_batch: the step's arguments (StepArgs per groom, ellipsoid
    // colliders, workgroup/brick prefix arrays) are laid out in one blob,
    // written into a pinned host slot and copied to `dev_blob` on the stream.
    // Two slots alternate; a slot is rewritten only after the copy that read
    // it has completed (slot_done).
    static constexpr int kPinSlots = 2;
    uint8_t* pin[kPinSlots] = {nullptr, nullptr};
    hipEvent_t slot_done[kPinSlots] = {nullptr, nullptr};
    bool slot_used[kPinSlots] = {false, false};
    int slot_next = 0;
    size_t pin_cap = 0;
    uint8_t* dev_blob = nullptr;
    size_t dev_cap = 0;
    // captured batch steps, keyed by the batch's shape (groom ids and kernel
    // variants); one executable graph per pinned slot
    bool graphs = false;
    hipStream_t cap_stream = nullptr;
    struct BatchGraph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec[kPinSlots] = {nullptr, nullptr};
        uint64_t last_use = 0;
    };
    std::vector<BatchGraph> graph_cache;
    uint64_t batch_calls = 0;
    uint64_t graph_captures = 0;
    hs_tile tile{hs::kWinCells, hs::kWinCells, 1, hs::kShapes[0].strands, hs::kShapes[0].points};
    int shape = 0;  // index of (tile.strands, tile.points) in hs::kShapes
};

namespace hsi {
// Device address of the fixed staging array of hs_read_strands_noalloc on
// `device` (looked up once per device, which must be current).
hipError_t noalloc_scratch(int device, void** out);
// Adds a built groom to the context; returns its id.
uint32_t ctx_adopt(hs_ctx* c, Groom&& g);
// Removes a groom from the context without freeing its block.
bool ctx_take(hs_ctx* c, uint32_t id, Groom* out);
}  // namespace hsi
