/*
 * hairsim.h - C ABI of the MI355X strand-physics pass.
 *
 * This library replaces the reference's Vulkan compute step: the one GLSL
 * kernel src/shaders/compute.comp:124-303 together with the host code that
 * binds its buffers and dispatches it (src/Renderer.cpp:2022-2077 records the
 * grid fill + vkCmdDispatch, src/Renderer.cpp:2309-2319 submits it each frame).
 *
 * Everything is plain C: opaque handles, plain pointers and sizes, status
 * codes.  No C++ exceptions cross this boundary and no HIP or torch type
 * appears in a signature.  All entry points return 0 (HS_OK) on success and a
 * negative hs_status on failure; hs_last_error() then returns a message.
 *
 * Threading: one context per host thread, or serialise calls on a context
 * yourself.  Steps are stream-ordered and asynchronous; hs_sync() and the
 * hs_read_* calls block.
 */
#ifndef HAIRSIM_H
#define HAIRSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
typedef enum hs_status {
    HS_OK = 0,
    HS_ERR_INVALID = -1,     /* bad argument (null pointer, size out of range) */
    HS_ERR_NO_DEVICE = -2,   /* no HIP device / wrong device ordinal          */
    HS_ERR_OUT_OF_MEMORY = -3,
    HS_ERR_HIP = -4,         /* a HIP runtime call failed                      */
    HS_ERR_NOT_FOUND = -5,   /* unknown groom / mesh id                        */
    HS_ERR_STATE = -6,       /* call not valid in the current state            */
    HS_ERR_BUDGET = -7       /* admission refused: HBM budget exhausted        */
} hs_status;

/* ------------------------------------------------------------------------ */
/* Boundary types: byte-for-byte the reference's buffer formats              */
/* ------------------------------------------------------------------------ */

/* Reference `Time` UBO, src/Scene.h:16-19 (8 B). */
typedef struct hs_time {
    float deltaTime;
    float totalTime;
} hs_time;

/* Reference `Collider`, src/Scene.h:23-39 (three column-major mat4, 192 B).
 * Collider 0 is the sphere (compute.comp:160-169: only transform[3].xyz, the
 * centre, is used); colliders 1.. are ellipsoids (compute.comp:170-179). */
typedef struct hs_collider {
    float transform[16];
    float inv[16];
    float invTrans[16];
} hs_collider;

/* Reference `GridCell`, src/Scene.h:42-49 / compute.comp:35-38 (16 B):
 * fixed-point (x1e6) velocity sum and density sum of one voxel. */
typedef struct hs_grid_cell {
    int32_t velocity[3];
    int32_t density;
} hs_grid_cell;

/* Reference `StrandDrawIndirect`, src/Strand.h:53-58 (16 B). */
typedef struct hs_draw_indirect {
    uint32_t vertexCount;
    uint32_t instanceCount;
    uint32_t firstVertex;
    uint32_t firstInstance;
} hs_draw_indirect;

/* The strand state itself crosses the boundary in the reference `Strand`
 * layout (src/Strand.h:11-16, compute.comp:44-48) with a run-time N:
 *     float strand[12*N] = { vec4 curvePoints[N]; vec4 curveVels[N];
 *                            vec4 correctionVecs[N]; }
 * i.e. 48*N bytes per strand, strands back to back. */
#define HS_STRAND_FLOATS(N) (12u * (uint32_t)(N))

#define HS_MAX_COLLIDERS 8u
#define HS_MAX_POINTS 256u

/* ------------------------------------------------------------------------ */
/* Groom (one strand buffer = one reference `Hair`)                          */
/* ------------------------------------------------------------------------ */

/* Keep the per-point correction vectors (compute.comp:201) in device state so
 * hs_read_strands can return them.  Off by default: they are per-step
 * temporaries of the shader (written at :201, consumed at :214) and storing
 * them costs 12 B per point per step. */
#define HS_GROOM_KEEP_CORRECTIONS 0x1u

typedef struct hs_groom_desc {
    uint32_t num_strands;     /* S >= 1                                       */
    uint32_t num_points;      /* N in [2, HS_MAX_POINTS] (NUM_CURVE_POINTS)   */
    /* Rest length of one segment.  When seg_length is NULL every strand uses
     * strand_length / (N - 1.0f), the reference's `radius`
     * (compute.comp:139-140, strandLength = 2.5).  Otherwise seg_length[s]
     * is strand s's segment rest length. */
    float strand_length;
    const float* seg_length;  /* optional, S floats                           */
    /* Optional per-strand wind noise: strand s feels an extra acceleration
     * wind_noise[s] * noise_dir (no reference counterpart: its wind is
     * commented out at compute.comp:151-152). */
    const float* wind_noise;  /* optional, S floats                           */
    float noise_dir[3];
    /* Hair-hair voxel grid (compute.comp:9-11, 204-206): G^3 cells of side h
     * from origin.  grid_dim == 0 disables the grid for this groom.  The
     * reference uses G = 64, h = 7/64, origin (-3,-2,-5). */
    uint32_t grid_dim;
    float grid_origin[3];
    float grid_h;
    uint32_t flags;           /* HS_GROOM_*                                   */
} hs_groom_desc;

/* ------------------------------------------------------------------------ */
/* Per-step dispatch parameters                                              */
/* ------------------------------------------------------------------------ */

/* Every constant compute.comp hard-codes becomes a field here; the values in
 * brackets are the reference's. */
typedef struct hs_step_params {
    float gravity[3];        /* (0,-9.8,0)              compute.comp:150       */
    float wind[3];           /* uniform wind acceleration, added to gravity    */
    float damping;           /* DAMPING 0.998           compute.comp:7         */
    float spring_k;          /* penalty stiffness 1900  compute.comp:163,173   */
    float sphere_radius;     /* collider 0 radius 1.0   compute.comp:161       */
    float vmax;              /* velocity clamp 10       compute.comp:198-200   */
    float friction;          /* hair-hair friction 0.08 compute.comp:296       */
    uint32_t num_colliders;  /* colliders used (NUM_COLLIDERS 6)               */
    uint32_t hair_hair;      /* 1: splat + grid friction; 0: skip the grid     */
    uint32_t mesh_id;        /* 0: none; else a mesh collider (hs_mesh_create) */
    uint32_t forcing;        /* wind forcing the shader carries commented out:
                              * 0 none (the reference), 1 compute.comp:150,
                              * 2 compute.comp:151 (fbm-modulated); sin/cos
                              * are one fixed binary32 algorithm, absolute
                              * error < 1e-7 for |x| <= 1e5                  */
} hs_step_params;

/* Fill `p` with the reference's constants (DAMPING 0.998, k 1900, sphere
 * radius 1.0, vmax 10, friction 0.08, gravity (0,-9.8,0), 6 colliders,
 * hair-hair on, no wind, no mesh, no forcing). */
void hs_step_params_reference(hs_step_params* p);

/* Fill `d` with the reference groom defaults for S strands of N points:
 * strand length 2.5, grid 64^3 with h = 7/64 at (-3,-2,-5). */
void hs_groom_desc_reference(hs_groom_desc* d, uint32_t num_strands, uint32_t num_points);

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct hs_ctx hs_ctx;

/* Replaces the Vulkan instance/device/queue bring-up for the compute path
 * (src/Instance.cpp:315, src/Device.cpp:19-21): bind HIP device `device` and
 * create the context's stream. */
int hs_create(hs_ctx** out, int device);
void hs_destroy(hs_ctx* ctx);
const char* hs_last_error(const hs_ctx* ctx);
/* Block until all work queued on the context has finished. */
int hs_sync(hs_ctx* ctx);
/* Run the context's work on a caller-owned hipStream_t (passed as void*), or
 * NULL to return to the context's own stream. */
int hs_set_stream(hs_ctx* ctx, void* hip_stream);
/* Device bytes currently held by the context (strand state, grids, meshes). */
int hs_memory_used(const hs_ctx* ctx, uint64_t* bytes);

/* ------------------------------------------------------------------------ */
/* Grooms                                                                    */
/* ------------------------------------------------------------------------ */

/* Replaces `Hair::Hair` uploading its strand SSBO and indirect-draw buffer
 * (src/Strand.cpp:149-191, BufferUtils::CreateBufferFromData at :188-190).
 * `strands_aos` holds S strands in the reference Strand layout (48*N B each);
 * it is copied, the caller keeps ownership. */
int hs_groom_create(hs_ctx* ctx, const hs_groom_desc* desc,
                    const float* strands_aos, uint32_t* groom_id);
/* Replaces `Hair::~Hair` (src/Strand.cpp:214-223). */
int hs_groom_destroy(hs_ctx* ctx, uint32_t groom_id);
/* Device bytes one groom of this shape needs (state + grids). */
uint64_t hs_groom_bytes(uint32_t num_strands, uint32_t num_points,
                        uint32_t grid_dim, uint32_t flags);

/* One physics step of one groom: the per-frame `Renderer::Frame` compute
 * submit (src/Renderer.cpp:2309-2319), i.e. grid clear (:2063) + dispatch of
 * compute.comp (:2070) with the Time UBO (compute.comp:20-23) and Collider UBO
 * (compute.comp:25-33) contents given here.  Asynchronous. */
int hs_step(hs_ctx* ctx, uint32_t groom_id, const hs_time* time,
            const hs_collider* colliders, uint32_t num_colliders,
            const hs_step_params* params);

/* One step of many grooms at once.  The reference steps each `Hair` with its
 * own vkCmdDispatch inside the one compute command buffer
 * (src/Renderer.cpp:2066-2071, where the descriptor index `4 + i` only works
 * for one groom); here all grooms' strand blocks go to one step launch and
 * all their voxel bricks to one reduce launch (per group of grooms that need
 * the same kernel variant).  Each groom's result is identical to stepping it
 * alone with hs_step.  A groom may appear at most once per call. */
typedef struct hs_batch_item {
    uint32_t groom_id;
    hs_time time;
    const hs_collider* colliders;
    uint32_t num_colliders;
    const hs_step_params* params;
} hs_batch_item;

int hs_step_batch(hs_ctx* ctx, const hs_batch_item* items, uint32_t num_items);

/* When on, hs_step_batch replays a captured HIP graph (argument copy, step
 * kernels, reduce kernels) for each batch shape it has seen: one graph launch
 * per step instead of one API call per kernel.  A shape is the list of groom
 * ids with their kernel variants; up to 16 shapes are kept.  Results are the
 * same bit for bit either way.  No reference counterpart (the reference
 * re-submits one pre-recorded command buffer, Renderer.cpp:2022-2077, 2317). */
int hs_set_batch_graphs(hs_ctx* ctx, int on);
/* hs_step_batch calls so far and the graphs captured for them. */
int hs_batch_stats(const hs_ctx* ctx, uint64_t* batch_calls, uint64_t* graph_captures);

/* Tile choices of the step kernels (no reference counterpart; the reference
 * fixes one strand per invocation in workgroups of 32, compute.comp:4,13, and
 * Renderer.cpp:2070).  Results are identical for every choice; only the speed
 * changes.  One lane steps one strand (64 strands per wavefront); the
 * workgroup takes `strands` strands and walks them root to tip `points`
 * points at a time.  (strands, points) is one of the built shapes: (128, 16),
 * (256, 8), (256, 12).  A change of shape takes effect on the next step. */
typedef struct hs_tile {
    uint32_t gather_cells;  /* voxels of the LDS window the friction gather stages, 64..2048 */
    uint32_t splat_cells;   /* voxels of one LDS splat window, 64..2048                   */
    uint32_t xcd_order;     /* 1: each XCD takes a contiguous run of workgroups           */
    uint32_t strands;       /* strands per workgroup                                      */
    uint32_t points;        /* points per segment of the root-to-tip walk                 */
} hs_tile;
int hs_set_tile(hs_ctx* ctx, const hs_tile* tile);
int hs_get_tile(const hs_ctx* ctx, hs_tile* tile);
/* Times every candidate tile (each built shape x gather and splat window sizes
 * x workgroup order) on copies of the batch's grooms (their state is left
 * untouched): each candidate starts from the grooms' present state and runs
 * one untimed and `steps` timed batched steps.  Keeps the fastest for later
 * steps (hs_step too).  best / best_ms may be NULL. */
int hs_autotune_batch(hs_ctx* ctx, const hs_batch_item* items, uint32_t num_items, uint32_t steps,
                      hs_tile* best, double* best_ms);

/* Read the groom back in the reference Strand layout (48*N B per strand) with
 * the velocities after this step's grid friction, plus the indirect draw
 * record {S, 1, 0, 0} the shader's count at compute.comp:302 produces.
 * Either output may be NULL.  Blocks. */
int hs_read_strands(hs_ctx* ctx, uint32_t groom_id, float* strands_aos,
                    hs_draw_indirect* indirect);
/* Render-facing export: the strand SSBO the reference renderer binds as its
 * vertex buffer and the indirect draw record it draws with (Strand.h:11-58,
 * Renderer.cpp:2240-2254, 2295), written into DEVICE memory: the groom in the
 * Strand layout (pending grid friction applied) into `strands_dev` (48*S*N
 * bytes) and, unless NULL, {S, 1, 0, 0} into `indirect_dev` (16 B).  Enqueued
 * on the library stream (hs_sync waits); nothing crosses the host link. */
int hs_export_strands(hs_ctx* ctx, uint32_t groom_id, float* strands_dev, uint32_t* indirect_dev);
/* The curve the hair tessellation shader draws through each strand
 * (hair.tese:34-80, `func`, one guide strand): M >= 2 samples per strand at
 * v = j/(M-1) on the piecewise cubic Bezier with control points
 * b1 = v1 + (1/3)((v2 - v0)/2), b2 = v2 - (1/3)((v3 - v1)/2), end neighbours
 * mirrored, de Casteljau with mix(a,b,t) = a(1-t) + bt; v = 1 is the tip
 * (where the shader would read one point past the strand).  Positions as
 * float4 (w = 1), S*M*16 bytes of DEVICE memory `out_dev`, caller strand
 * order.  Enqueued on the library stream. */
int hs_resample_strands(hs_ctx* ctx, uint32_t groom_id, uint32_t samples, float* out_dev);  /* 2 <= samples <= 2^20 */
/* hs_read_strands for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference.  Same outputs, bit for bit, but the device staging copy lives in
 * a fixed array of HS_NOALLOC_STAGING_BYTES per GPU that the library holds
 * until the process ends, so no call allocates or frees anything.  A groom of
 * more than HS_NOALLOC_STAGING_BYTES / 48 points (S*N) is refused with
 * HS_ERR_INVALID.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (from several threads or contexts) race, so serialise
 * them.  Blocks. */
#define HS_NOALLOC_STAGING_BYTES (64u << 20)
int hs_read_strands_noalloc(hs_ctx* ctx, uint32_t groom_id, float* strands_aos, hs_draw_indirect* indirect);
/* Overwrite the groom's state from a reference-layout buffer (checkpoint
 * restore; clears any pending grid friction). */
int hs_write_strands(hs_ctx* ctx, uint32_t groom_id, const float* strands_aos);
/* Read the voxel grid the last step splatted (compute.comp:217-252), G^3
 * cells in GridCell layout, index a + b*G + c*G*G.  Blocks. */
int hs_read_grid(hs_ctx* ctx, uint32_t groom_id, hs_grid_cell* cells);
/* Number of (cell, step) pairs since the groom was created whose GridCell
 * held a component of magnitude >= 2^30, i.e. within 2x of the int32 limit
 * of the shader's fixed-point sums (compute.comp:235-251, SCALE 1e6): the
 * format wraps past it, and it is kept as is.  Blocks. */
int hs_grid_near_limit(hs_ctx* ctx, uint32_t groom_id, uint64_t* cells);
/* Work of the last splatting step (no reference counterpart): the voxel
 * cells and boxes of the shared part of the partial-box pool (each
 * workgroup-segment's own box goes to a slot of its own and is not
 * counted) with the shared capacities, whether anything fell back to device
 * atomics (a box larger than any LDS window, or a full pool), and how many
 * workgroup-segments had a box larger than one window and were splatted
 * chunk by chunk by the workgroups that finished first.  Zeros before the
 * first such step.  Blocks. */
typedef struct hs_grid_work_stats {
    uint32_t pool_cells, pool_capacity, boxes, box_capacity, fallback, deferred;
} hs_grid_work_stats;
int hs_grid_work(hs_ctx* ctx, uint32_t groom_id, hs_grid_work_stats* out);

/* ------------------------------------------------------------------------ */
/* Triangle-mesh collider (no reference counterpart: compute.comp:155-184     */
/* handles a sphere and ellipsoids only).  Rule, per start-of-step point x:   */
/* q = the mesh point nearest to x (lowest triangle id on ties); contact when */
/* |x - q| < margin and x is inside (dot(x - q, n) < 0, n the angle-weighted  */
/* pseudo-normal of q's face/edge/vertex); the penalty k |x - q|              */
/* normalize(q - x) then counts as one more collider in the average of        */
/* compute.comp:182-184, after the UBO colliders.  Points deeper inside than  */
/* `margin` feel no mesh force.  A step uses the mesh hs_step_params.mesh_id  */
/* names (0: none).                                                           */
/* ------------------------------------------------------------------------ */

typedef struct hs_mesh_desc {
    /* triangle-grid cells per axis (each in [1, 1024]) over the mesh bounds
     * grown by margin.  The grid sets only the contact search's speed and
     * memory (48 B per listed cell-triangle pair), never its result; cells
     * about a quarter of the margin searched fastest on a head-sized mesh of
     * 20,480 triangles. */
    uint32_t grid_dim[3];
    float margin;          /* contact range, > 0 */
} hs_mesh_desc;

/* vertices: num_vertices x 3 floats; triangles: num_triangles x 3 indices,
 * counter-clockwise seen from outside; the mesh must be closed (every edge
 * shared by exactly two triangles) with finite coordinates.  Copies
 * everything. */
int hs_mesh_create(hs_ctx* ctx, const float* vertices, uint32_t num_vertices, const uint32_t* triangles,
                   uint32_t num_triangles, const hs_mesh_desc* desc, uint32_t* mesh_id);
int hs_mesh_destroy(hs_ctx* ctx, uint32_t mesh_id);

/* ------------------------------------------------------------------------ */
/* Kernel timing (no reference counterpart: the reference has no timers)     */
/* ------------------------------------------------------------------------ */

/* When on, every step launch is bracketed by HIP events recorded on the
 * stream the kernel runs on.  Enabling (or disabling) resets the totals. */
int hs_profile_enable(hs_ctx* ctx, int on);
/* Number of timed steps and their summed kernel time in ms.  Blocks. */
int hs_profile_read(hs_ctx* ctx, uint64_t* launches, double* total_ms);
/* The same total split into the strand step kernel and the voxel reduce
 * kernel (a step with hair-hair friction runs both).  Blocks. */
int hs_profile_breakdown(hs_ctx* ctx, double* step_kernel_ms, double* reduce_kernel_ms);
/* Summed time of the mesh pass (the per-point mesh contact kernel that
 * hs_step runs before the step kernel when a mesh collider is in use; it is
 * part of hs_profile_read's total, not of the step kernel's).  Blocks. */
int hs_profile_mesh(hs_ctx* ctx, double* mesh_kernel_ms);

/* ------------------------------------------------------------------------ */
/* Host-side collider helpers (src/Scene.h:28-38, src/Scene.cpp:110-136)      */
/* ------------------------------------------------------------------------ */

/* Collider(trans, rotDeg, scale): transform = T * Rz * Ry * Rx * S,
 * inv = inverse(transform), invTrans = transpose(inv). */
void hs_collider_from_trs(hs_collider* out, const float translation[3],
                          const float rotation_deg[3], const float scale[3]);
/* Scene::translateSphere: transform = transform * translate(t); inv and
 * invTrans recomputed. */
void hs_collider_translate(hs_collider* c, const float translation[3]);

/* ------------------------------------------------------------------------ */
/* Groom ingestion (host only; Strand.cpp:10-191, tinyobjloader 2.0)         */
/* ------------------------------------------------------------------------ */

/* The triangles of an OBJ file in file order, polygons triangulated the way
 * tinyobjloader 2.0 does with triangulate = true (ear clipping in the
 * dominant plane; a convex quad gives corners 0,1,2 then 0,2,3), numbers read
 * as it reads them: 9 floats of corner positions and 9 of corner normals per
 * triangle (GeneratePointsOnMesh's `positions` / `normals`, Strand.cpp:55-77).
 * Sets *num_triangles; fills the arrays when both are non-NULL and
 * `capacity` (triangles) suffices.  HS_ERR_NOT_FOUND: no such file;
 * HS_ERR_INVALID: a bad index or a corner without a normal. */
int hs_obj_triangles(const char* path, float* positions, float* normals, uint32_t capacity,
                     uint32_t* num_triangles);
/* Hair roots on those triangles (Strand.cpp:88-107): srand(seed) with MSVC's
 * rand() sequence (x = x*214013 + 2531011, bits 16..30; the reference seeds
 * 8), per strand a triangle rand() % T and barycentrics rand()/RAND_MAX folded
 * into it; the root normal is the triangle's first corner normal.  As in the
 * reference, rand() <= 32767 means only the first 32768 triangles can get
 * roots.  3 floats per strand in each output; 1 <= T <= 2^31 - 1. */
int hs_sample_roots(const float* positions, const float* normals, uint32_t num_triangles, uint32_t num_strands,
                    uint32_t seed, float* roots, float* root_normals);
/* Initial strands of Hair::Hair (Strand.cpp:157-172) in the Strand layout:
 * point j = root + j * (float)(length / (N - 1.0)) * (normal + (0.05, 5, -2)),
 * velocities (0, 0, -1), corrections 0.  48*N bytes per strand. */
int hs_initial_strands(const float* roots, const float* root_normals, uint32_t num_strands, uint32_t num_points,
                       float length, float* strands_aos);

/* ------------------------------------------------------------------------ */
/* Served simulation jobs (no reference counterpart: the reference steps one  */
/* groom per frame, main.cpp:257-285).  A server owns a context, an HBM arena */
/* of `hbm_budget` bytes and three threads: admission (uploads jobs in FIFO   */
/* order while the HBM budget allows; each admitted job goes to the step      */
/* loop with the fewest points), one or more step loops (each a context with  */
/* its own stream: each hs_step_batch takes the loop's resident jobs in       */
/* round-robin order up to max_points_per_step and steps each once; the loop  */
/* builds its next launch while the current one runs, with at most           */
/* launch_depth launches queued; the loops' launches run concurrently) and    */
/* completion (reads finished jobs back).  When a loop's jobs hold more       */
/* points than one launch, each advances once every few launches.             */
/* ------------------------------------------------------------------------ */

/* Wind and collider-0 centre at time t (seconds since the job's start),
 * linear between keyframes, held constant before the first and after the
 * last one. */
typedef struct hs_keyframe {
    float t;
    float wind[3];
    float sphere[3];
} hs_keyframe;

typedef struct hs_job {
    hs_groom_desc groom;          /* shape, rest lengths, noise, grid           */
    const float* strands_aos;     /* initial state, reference Strand layout     */
    uint32_t steps;               /* steps to run                               */
    float dt;                     /* step size; step k sees totalTime (k+1)*dt  */
    const hs_keyframe* keys;      /* wind/sphere trajectory (>= 1 key)          */
    uint32_t num_keys;
    hs_step_params params;        /* wind and collider 0 come from the keys;    */
                                  /* num_colliders is 0 or 1 (the sphere)        */
    float* result_aos;            /* receives the final state, 48*S*N bytes     */
                                  /* (NULL: the final state is dropped)          */
} hs_job;
/* strands_aos, groom.seg_length, groom.wind_noise, keys and result_aos are the
 * caller's and must stay valid until the job is reported done (hs_server_wait
 * / hs_server_drain): the admission thread reads the first three when it
 * uploads the job, the step loop reads the keys every step. */

#define HS_SERVER_MAX_STEP_LOOPS 8
typedef struct hs_server_config {
    int device;
    uint64_t hbm_budget;          /* device bytes of resident job state         */
    uint64_t max_points_per_step; /* points in one batched launch (latency cap) */
    int use_graphs;               /* replay captured graphs for repeated shapes */
    uint32_t step_loops;          /* step loops, each a context with its own    */
                                  /* stream; their launches run concurrently    */
                                  /* (0: 4, at most HS_SERVER_MAX_STEP_LOOPS)    */
    uint32_t launch_depth;        /* launches queued per loop, 1 or 2 (0: 1)    */
    hs_tile tile;                 /* step-kernel tile of every loop (all zero:  */
                                  /* 256 strands x 8 points, 2048-cell windows, */
                                  /* XCD order 0, the measured best for served  */
                                  /* job mixes)                                 */
} hs_server_config;

typedef struct hs_server_stats {
    uint64_t submitted, admitted, completed, cancelled;
    uint64_t launches;            /* batched step launches                      */
    uint64_t job_steps;           /* sum over launches of the jobs stepped      */
    uint64_t point_updates;       /* sum over launches of S*N stepped           */
    double p50_step_ms, p99_step_ms, max_step_ms;   /* per job-step: device     */
                                                    /* time of its launch       */
    double p50_wait_ms, p99_wait_ms;                /* submit -> first step     */
    double busy_s;                /* time with at least one job admitted        */
    uint64_t peak_resident_bytes;
    uint32_t peak_jobs_per_launch;
    double gpu_busy_s;            /* sum over launches of their device time     */
                                  /* (with several step loops launches overlap, */
                                  /* so it can exceed busy_s)                   */
    double p50_latency_ms, p99_latency_ms;  /* per job-step: host enqueue to    */
                                  /* completion seen by the host, the wait      */
                                  /* behind the launch queued before included   */
} hs_server_stats;

typedef struct hs_server hs_server;

int hs_server_create(const hs_server_config* cfg, hs_server** out);
/* Queues a job; returns at once.  HS_ERR_INVALID for anything hs_step_batch
 * would reject (steps 0, dt not finite and > 0, forcing > 2, more than one
 * collider, a mesh, a bad groom description), so a bad job never reaches a
 * launch shared with others.  HS_ERR_BUDGET if the job alone exceeds the
 * budget (its state, result staging and upload staging) or the per-step
 * point limit. */
int hs_server_submit(hs_server* srv, const hs_job* job, uint64_t* job_id);
/* Waits until the job is done (0), or timeout_s passes (HS_ERR_STATE);
 * a negative timeout waits without limit.  HS_ERR_NOT_FOUND for an unknown
 * or cancelled job. */
int hs_server_wait(hs_server* srv, uint64_t job_id, double timeout_s);
/* Drops queued jobs that are not admitted yet (their results stay unwritten). */
int hs_server_cancel_pending(hs_server* srv, uint64_t* cancelled);
/* Waits until every admitted job is done. */
int hs_server_drain(hs_server* srv);
int hs_server_get_stats(hs_server* srv, hs_server_stats* out);
/* Status and message of the first failure on a server thread (0 if none). */
int hs_server_error(hs_server* srv, const char** message);
/* Cancels what is queued, finishes what is admitted, frees everything. */
void hs_server_destroy(hs_server* srv);
/* The keyframe interpolation the server applies (for stepping a job alone). */
void hs_keyframes_eval(const hs_keyframe* keys, uint32_t num_keys, float t, float wind[3], float sphere[3]);

#ifdef __cplusplus
}
#endif
#endif /* HAIRSIM_H */
