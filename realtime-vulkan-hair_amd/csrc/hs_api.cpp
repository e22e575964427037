// hs_api.cpp - the C ABI (include/hairsim.h): context, grooms, step dispatch.
//
// This is the part of the reference's host that drives the compute pass:
// buffer creation (src/Strand.cpp:149-191, src/Scene.cpp:4-35), the recorded
// grid fill + dispatch (src/Renderer.cpp:2022-2077) and the per-frame submit
// (src/Renderer.cpp:2309-2319), re-expressed over one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hairsim.h"
#include "hs_kernels.h"

#include "hs_internal.h"
#include "hs_forcing.hpp"
#include "hs_mesh_build.h"

using namespace hsi;

namespace hsi {

ShapeDims shape_dims(int shape, uint32_t S, uint32_t N)
{
    const hs::TileShape t = hs::kShapes[shape];
    return ShapeDims{(S + t.strands - 1) / t.strands, (N - 1 + t.points - 1) / t.points};
}

Layout layout_of(uint32_t S, uint32_t N, uint32_t G)
{
    Layout l;
    l.Sp = (uint32_t)round_up(S, hs::kMaxTileStrands);
    l.nws_cap = 0;
    for (int k = 0; k < hs::kNumShapes; ++k) {
        const ShapeDims d = shape_dims(k, S, N);
        l.nws_cap = std::max(l.nws_cap, d.nwg * d.nseg);
    }
    l.pool_cap = 0;
    l.desc_cap = 0;
    if (G) {
        // each workgroup-segment's slot (its first box always fits), then a
        // shared part for further chunks: half a cell per point per step (at
        // least 64k cells) and three boxes per segment
        const uint64_t nws = l.nws_cap;
        l.pool_cap = nws * hs::kSlotCells + std::max<uint64_t>(65536, (uint64_t)S * (N - 1) / 2);
        l.desc_cap = nws + 3 * nws + 4096;
    }
    return l;
}

uint64_t carve(Groom& g, uint8_t* base)
{
    uint64_t off = 0;
    auto take = [&](auto*& ptr, uint64_t bytes) {
        using T = std::remove_reference_t<decltype(*ptr)>;
        ptr = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += round_up(bytes, 256);
    };
    const Layout& L = g.L;
    const uint64_t planes = 3ull * g.N * L.Sp * 4;
    take(g.pos, planes);
    take(g.vel, planes);
    if (g.flags & HS_GROOM_KEEP_CORRECTIONS)
        take(g.corr, planes);
    else
        g.corr = nullptr;
    take(g.seg, (uint64_t)L.Sp * 4);
    take(g.perm, (uint64_t)L.Sp * 4);
    take(g.cols, sizeof(hs::Collider) * hs::kMaxColliders);
    if (g.has_noise)
        take(g.noise, (uint64_t)L.Sp * 4);
    else
        g.noise = nullptr;
    g.gw = hs::GridWork{};
    if (g.G) {
        const uint64_t cells = grid_cells(g.G);
        take(g.gw.grid_i, cells * 16);
        take(g.gw.grid_f, cells * 16);
        take(g.gw.overflow, cells * 16);
        take(g.gw.pool, 2 * L.pool_cap * 16);
        take(g.gw.desc, 2 * L.desc_cap * hs::kDescInts * 4);
        take(g.gw.counters, 8 * 4);
        take(g.gw.gbox, (uint64_t)L.nws_cap * 6 * 4);
        take(g.gw.brick_cnt, bricks(g.G) * 2 * 4);
        take(g.gw.brick_list, bricks(g.G) * 2 * hs::kBrickList * 4);
        take(g.gw.brick_nz, bricks(g.G) * 4);  // zero: every brick of the zeroed grid is empty
        take(g.gw.near_limit, 8);
        take(g.gw.dlist, 2ull * L.nws_cap * 8);
        take(g.gw.dclaim, 2ull * L.nws_cap * 4);
        take(g.gw.dbox, (uint64_t)L.nws_cap * hs::kDBoxInts * 4);
        g.gw.pool_cap = (uint32_t)L.pool_cap;
        g.gw.desc_cap = (uint32_t)L.desc_cap;
        g.gw.nws = L.nws_cap;  // each launch sets its tile shape's count
        g.gw.nws_cap = L.nws_cap;
        g.gw.nbx = bricks_x(g.G);
        g.gw.nby = bricks_y(g.G);
        g.gw.nbz = bricks_z(g.G);
    }
    return off;
}

uint32_t ctx_adopt(hs_ctx* c, Groom&& g)
{
    c->bytes += g.bytes;
    const uint32_t gid = c->next_id++;
    c->grooms[gid] = std::move(g);
    return gid;
}

bool ctx_take(hs_ctx* c, uint32_t id, Groom* out)
{
    auto it = c->grooms.find(id);
    if (it == c->grooms.end()) return false;
    c->bytes -= it->second.bytes;
    *out = std::move(it->second);
    c->grooms.erase(it);
    return true;
}

void groom_release(Groom& g)
{
    if (g.block) {
        if (g.heap)
            g.heap->release(g.block);
        else
            (void)hipFree(g.block);
    }
    g = Groom();
}

}  // namespace hsi

namespace {

int fail(hs_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

int hip_fail(hs_ctx* c, hipError_t e, const char* what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(c, e == hipErrorOutOfMemory ? HS_ERR_OUT_OF_MEMORY : HS_ERR_HIP, m);
}

#define HS_TRY(ctx, expr)                                      \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

void free_groom(hs_ctx* c, Groom& g)
{
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->bytes -= g.bytes;
    groom_release(g);
}

Groom* find(hs_ctx* c, uint32_t id)
{
    auto it = c->grooms.find(id);
    return it == c->grooms.end() ? nullptr : &it->second;
}

int drain_profile(hs_ctx* c)
{
    for (size_t k = 0; k < c->events_used; ++k) {
        auto& ev = c->events[k];
        HS_TRY(c, hipEventSynchronize(ev.e[ev.reduce ? 2 : 1]));
        float ms = 0.0f;
        HS_TRY(c, hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
        c->prof_step_ms += ms;
        if (ev.mesh) {
            HS_TRY(c, hipEventElapsedTime(&ms, ev.e[3], ev.e[0]));
            c->prof_mesh_ms += ms;
        }
        if (ev.reduce) {
            HS_TRY(c, hipEventElapsedTime(&ms, ev.e[1], ev.e[2]));
            c->prof_reduce_ms += ms;
        }
        c->prof_launches += 1;
    }
    c->events_used = 0;
    return HS_OK;
}

void drop_graphs(hs_ctx* c)
{
    for (auto& g : c->graph_cache)
        for (int k = 0; k < hs_ctx::kPinSlots; ++k)
            if (g.exec[k]) (void)hipGraphExecDestroy(g.exec[k]);
    c->graph_cache.clear();
}

// Next set of timing events from the pool (created on first use).
int next_events(hs_ctx* c, hs_ctx::Ev** out)
{
    if (c->events_used == c->events.size()) {
        if (c->events.size() >= 4096) {
            int rc = drain_profile(c);
            if (rc != HS_OK) return rc;
        } else {
            hs_ctx::Ev ev{};
            for (int k = 0; k < 4; ++k) HS_TRY(c, hipEventCreate(&ev.e[k]));
            c->events.push_back(ev);
        }
    }
    *out = &c->events[c->events_used++];
    return HS_OK;
}

// Largest float T with sqrtf(T) <= vmax, so that for every non-NaN d,
// d > T exactly when sqrtf(d) > vmax (sqrtf is correctly rounded, hence
// monotonic).  Lets the kernel test |v| > vmax without a square root.
float sqrt_threshold(float vmax)
{
    if (std::isnan(vmax)) return vmax;                 // no d compares greater: never clamp
    if (vmax < 0.0f) return -INFINITY;                 // sqrt(d) >= 0 > vmax for all d
    if (std::isinf(vmax)) return INFINITY;
    float x = vmax * vmax;
    while (x > 0.0f && !(std::sqrt(x) <= vmax)) x = std::nextafter(x, 0.0f);
    for (;;) {
        const float nx = std::nextafter(x, INFINITY);
        if (!std::isfinite(nx) || !(std::sqrt(nx) <= vmax)) break;
        x = nx;
    }
    return x;
}

// Smallest float T with sqrtf(T) >= r, so that for every non-NaN d >= 0,
// d < T exactly when sqrtf(d) < r.  Lets the kernel test a sphere hit
// without a square root.
float sqrt_lower_threshold(float r)
{
    if (std::isnan(r)) return -INFINITY;   // sqrt(d) < NaN is never true
    if (!(r > 0.0f)) return 0.0f;          // sqrt(d) >= 0 >= r
    if (std::isinf(r)) return INFINITY;
    float x = r * r;
    if (std::isinf(x)) x = std::numeric_limits<float>::max();
    while (x > 0.0f && std::sqrt(std::nextafter(x, 0.0f)) >= r) x = std::nextafter(x, 0.0f);
    while (std::sqrt(x) < r) x = std::nextafter(x, INFINITY);
    return x;
}

// Physical strand order: a balanced k-d split of the roots into leaves of
// exactly kLeafStrands strands (only the last leaf may be short).  Nodes
// larger than kMaxTileStrands split at a multiple of kMaxTileStrands, so every
// aligned run of kMaxTileStrands slots is one subtree: a workgroup of any tile
// shape (whole leaves, aligned) holds strands that start close together, and
// the voxels they touch form a compact box; unlike a space-filling curve
// there is no jump between distant patches inside one workgroup.
std::vector<uint32_t> kd_perm(const float* aos, uint32_t S, uint32_t N)
{
    std::vector<float> root(3ull * S);
    for (uint32_t s = 0; s < S; ++s)
        for (int k = 0; k < 3; ++k) {
            const float v = aos[(size_t)s * 12 * N + k];
            root[3ull * s + k] = std::isnan(v) ? std::numeric_limits<float>::max() : v;
        }
    std::vector<uint32_t> idx(S);
    std::iota(idx.begin(), idx.end(), 0u);
    struct Range {
        uint32_t b, e;
    };
    std::vector<Range> stack{{0, S}};
    while (!stack.empty()) {
        const Range r = stack.back();
        stack.pop_back();
        const uint32_t n = r.e - r.b;
        if (n <= hs::kLeafStrands) continue;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = r.b; i < r.e; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], root[3ull * idx[i] + k]);
                hi[k] = std::max(hi[k], root[3ull * idx[i] + k]);
            }
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
        // split at the multiple of q nearest the median: the left part is
        // always whole leaves (whole largest workgroups above that size), the
        // remainder travels right
        const uint32_t q = n > hs::kMaxTileStrands ? hs::kMaxTileStrands : hs::kLeafStrands;
        uint32_t mid = (uint32_t)(((n / 2) + q / 2) / q) * q;
        mid = std::min(std::max(mid, q), n - 1);
        auto cmp = [&](uint32_t a, uint32_t b) {
            const float fa = root[3ull * a + ax], fb = root[3ull * b + ax];
            return fa < fb || (fa == fb && a < b);
        };
        std::nth_element(idx.begin() + r.b, idx.begin() + r.b + mid, idx.begin() + r.e, cmp);
        stack.push_back({r.b + mid, r.e});
        stack.push_back({r.b, r.b + mid});
    }
    return idx;
}

// Validate one step request and fill the kernel arguments for groom g.
// `stage`: when given, ellipsoid colliders are written there (the caller
// points a.cols at their device copy); otherwise they are uploaded to the
// groom's collider buffer on the stream.
int build_args(hs_ctx* c, Groom& g, const hs_time* t, const hs_collider* cols, uint32_t ncol,
               const hs_step_params* p, const char* who, hs::StepArgs* out, bool* gather, bool* splat,
               hs::Collider* stage = nullptr)
{
    const std::string w(who);
    if (!t || !p) return fail(c, HS_ERR_INVALID, w + ": null time or params");
    if (ncol > HS_MAX_COLLIDERS) return fail(c, HS_ERR_INVALID, w + ": more than 8 colliders");
    if (ncol > 0 && !cols) return fail(c, HS_ERR_INVALID, w + ": null colliders");
    if (p->num_colliders > ncol) return fail(c, HS_ERR_INVALID, w + ": params.num_colliders > colliders given");
    if (!(t->deltaTime > 0.0f) || !std::isfinite(t->deltaTime))
        return fail(c, HS_ERR_INVALID, w + ": deltaTime must be finite and > 0");
    if (p->forcing > 2) return fail(c, HS_ERR_INVALID, w + ": forcing must be 0, 1 or 2");
    const hsi::Mesh* mesh = nullptr;
    if (p->mesh_id != 0) {
        auto it = c->meshes.find(p->mesh_id);
        if (it == c->meshes.end()) return fail(c, HS_ERR_NOT_FOUND, w + ": unknown mesh");
        mesh = &it->second;
    }
    hs::StepArgs& a = *out;
    std::memset(&a, 0, sizeof(a));
    a.pos = g.pos;
    a.vel = g.vel;
    a.corr = g.corr;
    a.seg = g.seg;
    a.noise = g.noise;
    a.S = g.S;
    a.N = g.N;
    a.Sp = g.L.Sp;
    const ShapeDims sd = shape_dims(c->shape, g.S, g.N);
    a.nwg = sd.nwg;
    a.nseg = sd.nseg;
    a.gw = g.gw;
    a.gw.nws = sd.nwg * sd.nseg;
    a.gbox_ok = g.last_shape == c->shape ? 1u : 0u;
    a.G = g.G;
    a.parity = g.parity;
    a.tag = g.steps % 0x0fffffffu + 1u;  // 28 bits, never 0: the deferred list starts zero-filled
    a.ox = g.origin[0];
    a.oy = g.origin[1];
    a.oz = g.origin[2];
    a.inv_h = g.G ? 1.0f / g.h : 0.0f;
    a.dt = t->deltaTime;
    a.inv_dt = 1.0f / a.dt;
    a.dt2 = a.dt * a.dt;
    a.D = p->damping;
    a.k = p->spring_k;
    a.vmax = p->vmax;
    a.vmax2 = sqrt_threshold(p->vmax);
    a.sphere_r = p->sphere_radius;
    a.sphere_r2 = sqrt_lower_threshold(p->sphere_radius);
    a.fx = p->gravity[0] + p->wind[0];
    a.fy = p->gravity[1] + p->wind[1];
    a.fz = p->gravity[2] + p->wind[2];
    a.ndx = g.noise_dir[0];
    a.ndy = g.noise_dir[1];
    a.ndz = g.noise_dir[2];
    if (p->forcing) {
        // per-step constants of the forcing (fbm is uniform over the groom)
        const hs::Forcing fc = hs::forcing_at(p->forcing, t->totalTime);
        a.forcing = fc.mode;
        a.f_amp = fc.amp;
        a.f_a2 = fc.a2;
        a.f_t3 = fc.t3;
    }
    a.ncol = p->num_colliders;
    a.cols = g.cols;
    if (mesh) a.mesh = mesh->args;
    a.win_gather = c->tile.gather_cells;
    a.win_splat = c->tile.splat_cells;
    a.swz = c->tile.xcd_order;
    if (a.ncol > 0) {
        a.sphere_c[0] = cols[0].transform[12];
        a.sphere_c[1] = cols[0].transform[13];
        a.sphere_c[2] = cols[0].transform[14];
    }
    if (a.ncol > 1) {
        // ellipsoids are read from device memory; stream order keeps the
        // previous step's kernel ahead of this copy
        hs::Collider host[hs::kMaxColliders];
        hs::Collider* dst = stage ? stage : host;
        for (uint32_t j = 0; j < a.ncol; ++j) {
            std::memcpy(dst[j].m, cols[j].transform, 64);
            std::memcpy(dst[j].m + 16, cols[j].inv, 64);
            std::memcpy(dst[j].m + 32, cols[j].invTrans, 64);
        }
        if (!stage)
            HS_TRY(c, hipMemcpyAsync(g.cols, host, sizeof(hs::Collider) * a.ncol, hipMemcpyHostToDevice, c->stream));
    }
    *gather = g.pending;
    if (*gather) {
        a.mu_prev = g.pend_mu;
        a.omu_prev = g.pend_omu;
    }
    *splat = p->hair_hair && g.G > 0;
    return HS_OK;
}

// Host-side state after a launched step: pending friction, pool parity and
// the tile shape its boxes were written in.
void advance_state(Groom& g, bool splat, float friction, int shape)
{
    g.steps += 1;
    if (splat) {
        g.last_shape = shape;
        g.parity ^= 1u;
        g.pending = true;
        g.pend_mu = friction;
        g.pend_omu = 1.0f - friction;
    } else {
        g.pending = false;
    }
}

}  // namespace

namespace {

uint64_t align256(uint64_t x) { return round_up(x, 256); }

// Make room for a blob of `bytes` in the pinned slots and on the device.
// Growing drops the captured graphs (they hold the old addresses).
int ensure_blob(hs_ctx* c, size_t bytes)
{
    if (bytes <= c->pin_cap && bytes <= c->dev_cap) return HS_OK;
    const size_t cap = std::max<size_t>(bytes * 2, 64 * 1024);
    HS_TRY(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    for (int k = 0; k < hs_ctx::kPinSlots; ++k) {
        if (c->pin[k]) (void)hipHostFree(c->pin[k]);
        c->pin[k] = nullptr;
        c->slot_used[k] = false;
    }
    (void)hipFree(c->dev_blob);
    c->dev_blob = nullptr;
    c->pin_cap = c->dev_cap = 0;
    for (int k = 0; k < hs_ctx::kPinSlots; ++k) {
        HS_TRY(c, hipHostMalloc((void**)&c->pin[k], cap, hipHostMallocDefault));
        if (!c->slot_done[k]) HS_TRY(c, hipEventCreateWithFlags(&c->slot_done[k], hipEventDisableTiming));
    }
    HS_TRY(c, hipMalloc((void**)&c->dev_blob, cap));
    c->pin_cap = c->dev_cap = cap;
    return HS_OK;
}

// One kernel variant's share of a batch: which grooms, where their
// arguments and prefix arrays sit in the blob, and the grid sizes.
struct BatchLaunch {
    uint64_t args_off, wg_off, br_off;
    uint32_t ng, total_wg, total_br;
    uint8_t v;  // bit 0 gather, bit 1 splat, bit 2 ellipsoids, bit 3 mesh
};

int collider_mode(uint8_t v) { return ((v & 4) ? hs::kColEll : 0) | ((v & 8) ? hs::kColMesh : 0); }

// The batch's mesh pass: which StepArgs (indices into the blob's array) have
// a mesh, and the prefix sums of their blocks, both in the blob.
struct MeshPass {
    uint64_t idx_off = 0, blk_off = 0;
    uint32_t nm = 0, blocks = 0;
};

hipError_t launch_mesh_pass(const hs_ctx* c, const MeshPass& M, hipStream_t st)
{
    if (M.nm == 0) return hipSuccess;
    return hs::launch_mesh_force_batch(reinterpret_cast<const hs::StepArgs*>(c->dev_blob),
                                       reinterpret_cast<const uint32_t*>(c->dev_blob + M.idx_off),
                                       reinterpret_cast<const uint32_t*>(c->dev_blob + M.blk_off), M.nm, M.blocks, st);
}

// Room for `bytes` of mesh pass results (grow-only; the old buffer is freed
// once the stream is done with it).
int ensure_mesh_f(hs_ctx* c, size_t bytes)
{
    if (bytes <= c->mesh_f_cap) return HS_OK;
    HS_TRY(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->mesh_f);
    c->mesh_f = nullptr;
    c->mesh_f_cap = 0;
    HS_TRY(c, hipMalloc(&c->mesh_f, bytes));
    c->mesh_f_cap = bytes;
    return HS_OK;
}

hipError_t enqueue_batch(const hs_ctx* c, const std::vector<BatchLaunch>& launches, const MeshPass& mp, size_t bytes,
                         int slot, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(c->dev_blob, c->pin[slot], bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    e = launch_mesh_pass(c, mp, st);
    if (e != hipSuccess) return e;
    for (const BatchLaunch& L : launches) {
        e = hs::launch_step_batch(c->shape, reinterpret_cast<const hs::StepArgs*>(c->dev_blob + L.args_off),
                                  reinterpret_cast<const uint32_t*>(c->dev_blob + L.wg_off), L.ng, L.total_wg,
                                  (L.v & 1) != 0, (L.v & 2) != 0, collider_mode(L.v), c->tile.xcd_order, st);
        if (e != hipSuccess) return e;
    }
    for (const BatchLaunch& L : launches) {
        if (!(L.v & 2)) continue;
        e = hs::launch_reduce_batch(reinterpret_cast<const hs::StepArgs*>(c->dev_blob + L.args_off),
                                    reinterpret_cast<const uint32_t*>(c->dev_blob + L.br_off), L.ng, L.total_br, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// The executable graph for this batch shape and pinned slot, captured on
// first use.  At most kMaxGraphs shapes are kept (least recently used goes).
constexpr size_t kMaxGraphs = 16;

int batch_graph(hs_ctx* c, const std::vector<uint64_t>& key, const std::vector<BatchLaunch>& launches,
                const MeshPass& mp, size_t bytes, int slot, hipGraphExec_t* out)
{
    hs_ctx::BatchGraph* hit = nullptr;
    for (auto& g : c->graph_cache)
        if (g.key == key) hit = &g;
    if (!hit) {
        if (c->graph_cache.size() >= kMaxGraphs) {
            size_t lru = 0;
            for (size_t k = 1; k < c->graph_cache.size(); ++k)
                if (c->graph_cache[k].last_use < c->graph_cache[lru].last_use) lru = k;
            for (int k = 0; k < hs_ctx::kPinSlots; ++k)
                if (c->graph_cache[lru].exec[k]) (void)hipGraphExecDestroy(c->graph_cache[lru].exec[k]);
            c->graph_cache.erase(c->graph_cache.begin() + (long)lru);
        }
        c->graph_cache.emplace_back();
        hit = &c->graph_cache.back();
        hit->key = key;
    }
    hit->last_use = c->batch_calls;
    if (!hit->exec[slot]) {
        if (!c->cap_stream) HS_TRY(c, hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
        HS_TRY(c, hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeRelaxed));
        const hipError_t le = enqueue_batch(c, launches, mp, bytes, slot, c->cap_stream);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(c->cap_stream, &graph);
        if (le != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return hip_fail(c, le, "hs_step_batch: capture");
        }
        if (ce != hipSuccess) return hip_fail(c, ce, "hipStreamEndCapture");
        const hipError_t ie = hipGraphInstantiate(&hit->exec[slot], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            hit->exec[slot] = nullptr;
            return hip_fail(c, ie, "hipGraphInstantiate");
        }
        c->graph_captures += 1;
    }
    *out = hit->exec[slot];
    return HS_OK;
}

}  // namespace

extern "C" {


void hs_step_params_reference(hs_step_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->gravity[1] = -9.8f;   // compute.comp:150
    p->damping = 0.998f;     // compute.comp:7
    p->spring_k = 1900.0f;   // compute.comp:163,173
    p->sphere_radius = 1.0f; // compute.comp:161
    p->vmax = 10.0f;         // compute.comp:198
    p->friction = 0.08f;     // compute.comp:296
    p->num_colliders = 6;    // compute.comp:8
    p->hair_hair = 1;
    p->mesh_id = 0;
}

void hs_groom_desc_reference(hs_groom_desc* d, uint32_t S, uint32_t N)
{
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->num_strands = S;
    d->num_points = N;
    d->strand_length = 2.5f;           // compute.comp:139
    d->grid_dim = 64;                  // compute.comp:9
    d->grid_h = 7.0f / 64.0f;          // compute.comp:205 (GRID_HEIGHT / GRID_DIM)
    d->grid_origin[0] = -3.0f;         // compute.comp:206
    d->grid_origin[1] = -2.0f;
    d->grid_origin[2] = -5.0f;
    d->noise_dir[0] = 1.0f;
}

int hs_create(hs_ctx** out, int device)
{
    if (!out) return HS_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return HS_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return HS_ERR_NO_DEVICE;
    hs_ctx* c = new (std::nothrow) hs_ctx();
    if (!c) return HS_ERR_OUT_OF_MEMORY;
    c->device = device;
    void* staging = nullptr;  // the fixed staging array of hs_read_strands_noalloc, looked up once per device
    if (hipSetDevice(device) != hipSuccess || hsi::noalloc_scratch(device, &staging) != hipSuccess ||
        hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return HS_ERR_HIP;
    }
    c->stream = c->own;
    *out = c;
    return HS_OK;
}

void hs_destroy(hs_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->grooms) free_groom(c, kv.second);
    c->grooms.clear();
    for (auto& kv : c->meshes) (void)hipFree(kv.second.block);
    c->meshes.clear();
    for (auto& ev : c->events)
        for (int k = 0; k < 4; ++k) (void)hipEventDestroy(ev.e[k]);
    (void)hipFree(c->mesh_f);
    drop_graphs(c);
    for (int k = 0; k < hs_ctx::kPinSlots; ++k) {
        if (c->pin[k]) (void)hipHostFree(c->pin[k]);
        if (c->slot_done[k]) (void)hipEventDestroy(c->slot_done[k]);
    }
    (void)hipFree(c->dev_blob);
    if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

const char* hs_last_error(const hs_ctx* c) { return c ? c->err.c_str() : "null context"; }

int hs_sync(hs_ctx* c)
{
    if (!c) return HS_ERR_INVALID;
    HS_TRY(c, hipSetDevice(c->device));
    HS_TRY(c, hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_set_stream(hs_ctx* c, void* s)
{
    if (!c) return HS_ERR_INVALID;
    HS_TRY(c, hipStreamSynchronize(c->stream));
    c->stream = s ? reinterpret_cast<hipStream_t>(s) : c->own;
    return HS_OK;
}

int hs_memory_used(const hs_ctx* c, uint64_t* bytes)
{
    if (!c || !bytes) return HS_ERR_INVALID;
    *bytes = c->bytes;
    return HS_OK;
}

uint64_t hs_groom_bytes(uint32_t S, uint32_t N, uint32_t G, uint32_t flags)
{
    if (S == 0 || N < 2) return 0;
    Groom g;
    g.S = S;
    g.N = N;
    g.G = G;
    g.flags = flags;
    g.has_noise = true;
    g.L = layout_of(S, N, G);
    return carve(g, nullptr);
}

}  // extern "C"

namespace hsi {

int check_groom_desc(const hs_groom_desc* d, const float* aos, std::string* why)
{
    auto bad = [&](const char* m) {
        *why = m;
        return HS_ERR_INVALID;
    };
    if (!d || !aos) return bad("null argument");
    if (d->num_strands == 0) return bad("num_strands must be >= 1");
    if (d->num_points < 2 || d->num_points > HS_MAX_POINTS) return bad("num_points must be in [2, 256]");
    if (d->num_strands > (1u << 26)) return bad("too many strands");
    // kernels address the [3][N][Sp] state with 32-bit byte offsets: 12*N*Sp < 2^32
    if ((uint64_t)d->num_points * round_up(d->num_strands, hs::kMaxTileStrands) * 12 >= (1ull << 32))
        return bad("more than 357M points in one groom");
    if (d->grid_dim > 1024) return bad("grid_dim > 1024");
    if (d->grid_dim > 0 && !(d->grid_h > 0.0f && std::isfinite(d->grid_h)))
        return bad("grid_h must be finite and > 0");
    if (!d->seg_length && !(d->strand_length > 0.0f)) return bad("strand_length must be > 0");
    return HS_OK;
}

int groom_build(const hs_groom_desc* d, const float* aos, hipStream_t st, DevHeap* heap, Groom* out,
                std::string* why)
{
    Groom g;
    g.S = d->num_strands;
    g.N = d->num_points;
    g.G = d->grid_dim;
    g.flags = d->flags;
    g.h = d->grid_h;
    g.has_noise = d->wind_noise != nullptr;
    for (int k = 0; k < 3; ++k) {
        g.origin[k] = d->grid_origin[k];
        g.noise_dir[k] = d->noise_dir[k];
    }
    g.L = layout_of(g.S, g.N, g.G);
    const Layout& L = g.L;
    g.bytes = carve(g, nullptr);
    const size_t aos_bytes = (size_t)g.S * g.N * 48;
    g.heap = heap;
    void* staging = nullptr;
    auto alloc = [&](void** p, size_t bytes) { return heap ? heap->alloc(p, bytes) : hipMalloc(p, bytes); };
    auto cleanup = [&](hipError_t e, const char* what) {
        (void)hipStreamSynchronize(st);
        if (staging) {
            if (heap)
                heap->release(staging);
            else
                (void)hipFree(staging);
        }
        groom_release(g);
        *why = std::string(what) + ": " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? HS_ERR_OUT_OF_MEMORY : HS_ERR_HIP;
    };
#define HS_CHECK(expr)                                 \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return cleanup(e_, #expr); \
    } while (0)
    HS_CHECK(alloc(&g.block, g.bytes));
    carve(g, static_cast<uint8_t*>(g.block));
    HS_CHECK(hipMemsetAsync(g.block, 0, g.bytes, st));
    HS_CHECK(alloc(&staging, aos_bytes));

    // physical order: k-d leaves of the roots, one per workgroup
    const std::vector<uint32_t> perm = kd_perm(aos, g.S, g.N);
    std::vector<float> seg(L.Sp, 0.0f), nz;
    // radius = strandLength / (NUM_CURVE_POINTS - 1.0)  (compute.comp:139-140)
    const float r_uniform = d->seg_length ? 0.0f : d->strand_length / ((float)g.N - 1.0f);
    for (uint32_t p = 0; p < g.S; ++p) seg[p] = d->seg_length ? d->seg_length[perm[p]] : r_uniform;
    if (d->wind_noise) {
        nz.assign(L.Sp, 0.0f);
        for (uint32_t p = 0; p < g.S; ++p) nz[p] = d->wind_noise[perm[p]];
    }
    HS_CHECK(hipMemcpyAsync(g.seg, seg.data(), (size_t)L.Sp * 4, hipMemcpyHostToDevice, st));
    HS_CHECK(hipMemcpyAsync(g.perm, perm.data(), (size_t)g.S * 4, hipMemcpyHostToDevice, st));
    if (d->wind_noise) HS_CHECK(hipMemcpyAsync(g.noise, nz.data(), (size_t)L.Sp * 4, hipMemcpyHostToDevice, st));
    HS_CHECK(hipMemcpyAsync(staging, aos, aos_bytes, hipMemcpyHostToDevice, st));
    HS_CHECK(hs::launch_pack(static_cast<const float*>(staging), g.perm, g.S, g.N, L.Sp, g.pos, g.vel, g.corr, st));
    HS_CHECK(hipStreamSynchronize(st));
#undef HS_CHECK
    if (heap)
        heap->release(staging);
    else
        (void)hipFree(staging);
    *out = std::move(g);
    return HS_OK;
}

hipError_t groom_unpack(const Groom& g, float* aos_dev, hipStream_t st)
{
    hs::UnpackGrid ug;
    const hs::UnpackGrid* pug = nullptr;
    if (g.pending) {
        ug.grid = g.gw.grid_f;
        ug.G = g.G;
        ug.ox = g.origin[0];
        ug.oy = g.origin[1];
        ug.oz = g.origin[2];
        ug.inv_h = 1.0f / g.h;
        ug.mu = g.pend_mu;
        ug.omu = g.pend_omu;
        pug = &ug;
    }
    return hs::launch_unpack(aos_dev, g.perm, g.S, g.N, g.L.Sp, g.pos, g.vel, g.corr,
                             pug, st);
}

}  // namespace hsi

extern "C" {

int hs_groom_create(hs_ctx* c, const hs_groom_desc* d, const float* aos, uint32_t* id)
{
    if (!c) return HS_ERR_INVALID;
    if (!id) return fail(c, HS_ERR_INVALID, "hs_groom_create: null argument");
    std::string why;
    int rc = check_groom_desc(d, aos, &why);
    if (rc != HS_OK) return fail(c, rc, "hs_groom_create: " + why);
    HS_TRY(c, hipSetDevice(c->device));
    Groom g;
    rc = groom_build(d, aos, c->stream, nullptr, &g, &why);
    if (rc != HS_OK) return fail(c, rc, "hs_groom_create: " + why);
    *id = ctx_adopt(c, std::move(g));
    return HS_OK;
}

int hs_mesh_create(hs_ctx* c, const float* v, uint32_t nv, const uint32_t* t, uint32_t nt, const hs_mesh_desc* d,
                   uint32_t* id)
{
    if (!c) return HS_ERR_INVALID;
    if (!v || !t || !d || !id) return fail(c, HS_ERR_INVALID, "hs_mesh_create: null argument");
    if (!(d->margin > 0.0f) || !std::isfinite(d->margin))
        return fail(c, HS_ERR_INVALID, "hs_mesh_create: margin must be finite and > 0");
    for (int k = 0; k < 3; ++k)
        if (d->grid_dim[k] == 0 || d->grid_dim[k] > 1024)
            return fail(c, HS_ERR_INVALID, "hs_mesh_create: grid_dim must be in [1, 1024]");
    hsmesh::Built b;
    if (!hsmesh::build(v, nv, t, nt, d->grid_dim, d->margin, &b))
        return fail(c, HS_ERR_INVALID,
                    "hs_mesh_create: mesh is not closed (or an index is out of range, or a vertex is not finite)");
    HS_TRY(c, hipSetDevice(c->device));
    const uint64_t tri_bytes = round_up(b.tris.size() * 4, 256);
    const uint64_t start_bytes = round_up(b.cell_start.size() * 4, 256);
    const uint64_t rec_bytes = round_up(std::max<size_t>(b.cell_recs.size(), hsmesh::kRecFloats) * 4, 256);
    hsi::Mesh m;
    m.bytes = tri_bytes + start_bytes + rec_bytes;
    m.ntris = nt;
    HS_TRY(c, hipMalloc(&m.block, m.bytes));
    uint8_t* base = static_cast<uint8_t*>(m.block);
    uint8_t* starts = base + tri_bytes;
    uint8_t* recs = starts + start_bytes;
    hipError_t e = hipMemcpyAsync(base, b.tris.data(), b.tris.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(starts, b.cell_start.data(), b.cell_start.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && !b.cell_recs.empty())
        e = hipMemcpyAsync(recs, b.cell_recs.data(), b.cell_recs.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(m.block);
        return hip_fail(c, e, "hs_mesh_create: upload");
    }
    hs::MeshArgs& a = m.args;
    a.tris = reinterpret_cast<const float4*>(base);
    a.cell_start = reinterpret_cast<const uint32_t*>(starts);
    a.cell_recs = reinterpret_cast<const float4*>(recs);
    a.half_diag = b.half_diag;
    a.ox = b.origin[0];
    a.oy = b.origin[1];
    a.oz = b.origin[2];
    a.icx = b.inv_cell[0];
    a.icy = b.inv_cell[1];
    a.icz = b.inv_cell[2];
    a.dx = b.dim[0];
    a.dy = b.dim[1];
    a.dz = b.dim[2];
    a.margin2 = d->margin * d->margin;
    const uint32_t mid = c->next_mesh++;
    c->meshes[mid] = m;
    c->bytes += m.bytes;
    *id = mid;
    return HS_OK;
}

int hs_mesh_destroy(hs_ctx* c, uint32_t id)
{
    if (!c) return HS_ERR_INVALID;
    auto it = c->meshes.find(id);
    if (it == c->meshes.end()) return fail(c, HS_ERR_NOT_FOUND, "hs_mesh_destroy: unknown mesh");
    HS_TRY(c, hipSetDevice(c->device));
    HS_TRY(c, hipStreamSynchronize(c->stream));
    (void)hipFree(it->second.block);
    c->bytes -= it->second.bytes;
    c->meshes.erase(it);
    return HS_OK;
}

int hs_groom_destroy(hs_ctx* c, uint32_t id)
{
    if (!c) return HS_ERR_INVALID;
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_groom_destroy: unknown groom");
    free_groom(c, *g);
    c->grooms.erase(id);
    return HS_OK;
}

int hs_step(hs_ctx* c, uint32_t id, const hs_time* t, const hs_collider* cols, uint32_t ncol,
            const hs_step_params* p)
{
    if (!c) return HS_ERR_INVALID;
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_step: unknown groom");
    HS_TRY(c, hipSetDevice(c->device));
    hs::StepArgs a;
    bool gather = false, splat = false;
    int rc = build_args(c, *g, t, cols, ncol, p, "hs_step", &a, &gather, &splat);
    if (rc != HS_OK) return rc;

    const bool mesh = a.mesh.tris != nullptr;
    if (mesh) {
        // the mesh pass writes every point's contact; the step kernel reads it
        rc = ensure_mesh_f(c, (size_t)g->N * g->L.Sp * sizeof(float4));
        if (rc != HS_OK) return rc;
        a.mesh_f = static_cast<const float4*>(c->mesh_f);
    }
    hs_ctx::Ev* ev = nullptr;
    if (c->profile) {
        rc = next_events(c, &ev);
        if (rc != HS_OK) return rc;
        ev->reduce = splat;
        ev->mesh = mesh;
        if (mesh) HS_TRY(c, hipEventRecord(ev->e[3], c->stream));
    }
    if (mesh) HS_TRY(c, hs::launch_mesh_force(a, c->stream));
    if (ev) HS_TRY(c, hipEventRecord(ev->e[0], c->stream));
    HS_TRY(c, hs::launch_step(c->shape, a, gather, splat, c->stream));
    if (ev) HS_TRY(c, hipEventRecord(ev->e[1], c->stream));
    if (splat) {
        HS_TRY(c, hs::launch_reduce(g->gw, g->G, g->parity, c->stream));
        if (ev) HS_TRY(c, hipEventRecord(ev->e[2], c->stream));
    }
    advance_state(*g, splat, p->friction, c->shape);
    return HS_OK;
}


int hs_step_batch(hs_ctx* c, const hs_batch_item* items, uint32_t n)
{
    if (!c) return HS_ERR_INVALID;
    if (n == 0) return HS_OK;
    if (!items) return fail(c, HS_ERR_INVALID, "hs_step_batch: null items");
    HS_TRY(c, hipSetDevice(c->device));
    // validate everything before touching any state
    std::vector<Groom*> gs(n);
    for (uint32_t k = 0; k < n; ++k) {
        gs[k] = find(c, items[k].groom_id);
        if (!gs[k]) return fail(c, HS_ERR_NOT_FOUND, "hs_step_batch: unknown groom");
        if (!items[k].params) return fail(c, HS_ERR_INVALID, "hs_step_batch: null params");
    }
    {
        std::vector<uint32_t> ids(n);
        for (uint32_t k = 0; k < n; ++k) ids[k] = items[k].groom_id;
        std::sort(ids.begin(), ids.end());
        if (std::adjacent_find(ids.begin(), ids.end()) != ids.end())
            return fail(c, HS_ERR_INVALID, "hs_step_batch: groom listed twice");
    }
    std::vector<hs::StepArgs> args(n);
    std::vector<uint8_t> variant(n);
    std::vector<hs::Collider> ell;           // ellipsoid colliders of every groom that has them
    std::vector<uint32_t> ell_at(n, 0);
    for (uint32_t k = 0; k < n; ++k) {
        bool gather = false, splat = false;
        hs::Collider stage[hs::kMaxColliders];
        int rc = build_args(c, *gs[k], &items[k].time, items[k].colliders, items[k].num_colliders, items[k].params,
                            "hs_step_batch", &args[k], &gather, &splat, stage);
        if (rc != HS_OK) return rc;
        const bool e = args[k].ncol > 1;
        variant[k] = (uint8_t)((gather ? 1 : 0) | (splat ? 2 : 0) | (e ? 4 : 0) | (args[k].mesh.tris ? 8 : 0));
        if (e) {
            ell_at[k] = (uint32_t)ell.size();
            ell.insert(ell.end(), stage, stage + args[k].ncol);
        }
    }
    // blob: [StepArgs of every group][ellipsoids][prefix arrays of every group]
    std::vector<std::vector<uint32_t>> groups(16);
    for (uint32_t k = 0; k < n; ++k) groups[variant[k]].push_back(k);
    std::vector<uint32_t> order;
    std::vector<uint32_t> idx;
    std::vector<BatchLaunch> launches;
    std::vector<uint64_t> key;
    for (uint8_t v = 0; v < 16; ++v) {
        if (groups[v].empty()) continue;
        BatchLaunch L{order.size() * sizeof(hs::StepArgs), 0, 0, (uint32_t)groups[v].size(), 0, 0, v};
        std::vector<uint32_t> wg{0}, br{0};
        key.push_back(((uint64_t)v << 32) | groups[v].size());
        for (uint32_t k : groups[v]) {
            order.push_back(k);
            key.push_back(((uint64_t)items[k].groom_id << 8) | args[k].ncol);
            wg.push_back(wg.back() + args[k].nwg);
            br.push_back(br.back() + ((v & 2) ? (uint32_t)bricks(args[k].G) : 0));
        }
        L.total_wg = wg.back();
        L.total_br = br.back();
        L.wg_off = idx.size();
        idx.insert(idx.end(), wg.begin(), wg.end());
        L.br_off = idx.size();
        idx.insert(idx.end(), br.begin(), br.end());
        launches.push_back(L);
    }
    // mesh pass: the grooms with a mesh, their blocks, and their slices of
    // the shared result buffer
    MeshPass mp;
    size_t mesh_f_bytes = 0;
    std::vector<size_t> mesh_at(n, 0);
    {
        std::vector<uint32_t> midx, mblk{0};
        for (size_t j = 0; j < order.size(); ++j) {
            const hs::StepArgs& a = args[order[j]];
            if (!a.mesh.tris) continue;
            const uint64_t plane = (uint64_t)a.N * a.Sp;
            midx.push_back((uint32_t)j);
            mblk.push_back(mblk.back() + (uint32_t)((plane + 255) / 256));
            mesh_at[order[j]] = mesh_f_bytes;
            mesh_f_bytes += plane * sizeof(float4);
        }
        if (!midx.empty()) {
            mp.nm = (uint32_t)midx.size();
            mp.blocks = mblk.back();
            mp.idx_off = idx.size();
            idx.insert(idx.end(), midx.begin(), midx.end());
            mp.blk_off = idx.size();
            idx.insert(idx.end(), mblk.begin(), mblk.end());
        }
    }
    const uint64_t args_bytes = align256(order.size() * sizeof(hs::StepArgs));
    const uint64_t ell_bytes = align256(ell.size() * sizeof(hs::Collider));
    const uint64_t bytes = args_bytes + ell_bytes + idx.size() * sizeof(uint32_t);
    for (BatchLaunch& L : launches) {
        L.wg_off = args_bytes + ell_bytes + L.wg_off * sizeof(uint32_t);
        L.br_off = args_bytes + ell_bytes + L.br_off * sizeof(uint32_t);
    }
    mp.idx_off = args_bytes + ell_bytes + mp.idx_off * sizeof(uint32_t);
    mp.blk_off = args_bytes + ell_bytes + mp.blk_off * sizeof(uint32_t);
    key.push_back(bytes);
    key.push_back(c->tile.xcd_order);
    key.push_back((uint64_t)c->shape);  // the captured kernels are the tile shape's
    key.push_back(mp.nm);
    int rc = ensure_blob(c, bytes);
    if (rc != HS_OK) return rc;
    if (mp.nm) {
        rc = ensure_mesh_f(c, mesh_f_bytes);
        if (rc != HS_OK) return rc;
        for (uint32_t k = 0; k < n; ++k)
            if (args[k].mesh.tris)
                args[k].mesh_f = reinterpret_cast<const float4*>(static_cast<uint8_t*>(c->mesh_f) + mesh_at[k]);
    }
    const int slot = c->slot_next;
    c->slot_next = (slot + 1) % hs_ctx::kPinSlots;
    if (c->slot_used[slot]) HS_TRY(c, hipEventSynchronize(c->slot_done[slot]));
    uint8_t* h = c->pin[slot];
    hs::Collider* ell_dev = reinterpret_cast<hs::Collider*>(c->dev_blob + args_bytes);
    for (size_t j = 0; j < order.size(); ++j) {
        hs::StepArgs a = args[order[j]];
        if (a.ncol > 1) a.cols = ell_dev + ell_at[order[j]];
        std::memcpy(h + j * sizeof(hs::StepArgs), &a, sizeof(a));
    }
    if (!ell.empty()) std::memcpy(h + args_bytes, ell.data(), ell.size() * sizeof(hs::Collider));
    std::memcpy(h + args_bytes + ell_bytes, idx.data(), idx.size() * sizeof(uint32_t));

    hipGraphExec_t exec = nullptr;
    if (c->graphs) {
        rc = batch_graph(c, key, launches, mp, bytes, slot, &exec);
        if (rc != HS_OK) return rc;
    }
    hs_ctx::Ev* ev = nullptr;
    bool any_splat = false;
    for (uint32_t k = 0; k < n; ++k) any_splat |= (variant[k] & 2) != 0;
    if (c->profile) {
        rc = next_events(c, &ev);
        if (rc != HS_OK) return rc;
        ev->reduce = false;
        ev->mesh = false;
    }
    if (exec) {
        // one graph: the mesh pass's time is counted with the step kernels
        if (ev) HS_TRY(c, hipEventRecord(ev->e[0], c->stream));
        HS_TRY(c, hipGraphLaunch(exec, c->stream));
        if (ev) HS_TRY(c, hipEventRecord(ev->e[1], c->stream));
    } else if (ev) {
        // eager launches: split the timing into step kernels and reduce kernels
        HS_TRY(c, hipMemcpyAsync(c->dev_blob, c->pin[slot], bytes, hipMemcpyHostToDevice, c->stream));
        if (mp.nm) {
            ev->mesh = true;
            HS_TRY(c, hipEventRecord(ev->e[3], c->stream));
            HS_TRY(c, launch_mesh_pass(c, mp, c->stream));
        }
        HS_TRY(c, hipEventRecord(ev->e[0], c->stream));
        for (const BatchLaunch& L : launches)
            HS_TRY(c, hs::launch_step_batch(c->shape, reinterpret_cast<const hs::StepArgs*>(c->dev_blob + L.args_off),
                                            reinterpret_cast<const uint32_t*>(c->dev_blob + L.wg_off), L.ng,
                                            L.total_wg, (L.v & 1) != 0, (L.v & 2) != 0, collider_mode(L.v),
                                            c->tile.xcd_order, c->stream));
        HS_TRY(c, hipEventRecord(ev->e[1], c->stream));
        for (const BatchLaunch& L : launches)
            if (L.v & 2)
                HS_TRY(c, hs::launch_reduce_batch(reinterpret_cast<const hs::StepArgs*>(c->dev_blob + L.args_off),
                                                  reinterpret_cast<const uint32_t*>(c->dev_blob + L.br_off), L.ng,
                                                  L.total_br, c->stream));
        if (any_splat) {
            ev->reduce = true;
            HS_TRY(c, hipEventRecord(ev->e[2], c->stream));
        }
    } else {
        HS_TRY(c, enqueue_batch(c, launches, mp, bytes, slot, c->stream));
    }
    HS_TRY(c, hipEventRecord(c->slot_done[slot], c->stream));
    c->slot_used[slot] = true;
    c->batch_calls += 1;
    for (uint32_t k = 0; k < n; ++k)
        advance_state(*gs[k], (variant[k] & 2) != 0, items[k].params->friction, c->shape);
    return HS_OK;
}

int hs_set_batch_graphs(hs_ctx* c, int on)
{
    if (!c) return HS_ERR_INVALID;
    HS_TRY(c, hipSetDevice(c->device));
    if (!on) {
        HS_TRY(c, hipStreamSynchronize(c->stream));
        drop_graphs(c);
    }
    c->graphs = on != 0;
    return HS_OK;
}

int hs_set_tile(hs_ctx* c, const hs_tile* t)
{
    if (!c) return HS_ERR_INVALID;
    if (!t) return fail(c, HS_ERR_INVALID, "hs_set_tile: null tile");
    if (t->gather_cells < 64 || t->gather_cells > hs::kWinCells || t->splat_cells < 64 ||
        t->splat_cells > hs::kWinCells || t->xcd_order > 1)
        return fail(c, HS_ERR_INVALID, "hs_set_tile: window sizes must be in [64, 2048], xcd_order 0 or 1");
    int shape = -1;
    for (int k = 0; k < hs::kNumShapes; ++k)
        if (hs::kShapes[k].strands == t->strands && hs::kShapes[k].points == t->points) shape = k;
    if (shape < 0)
        return fail(c, HS_ERR_INVALID,
                    "hs_set_tile: (strands, points) must be one of the built shapes (128, 16), (256, 8), (256, 12)");
    c->tile = *t;
    c->shape = shape;
    return HS_OK;
}

int hs_get_tile(const hs_ctx* c, hs_tile* t)
{
    if (!c || !t) return HS_ERR_INVALID;
    *t = c->tile;
    return HS_OK;
}

int hs_autotune_batch(hs_ctx* c, const hs_batch_item* items, uint32_t n, uint32_t steps, hs_tile* best,
                      double* best_ms)
{
    if (!c) return HS_ERR_INVALID;
    if (!items || n == 0 || steps == 0) return fail(c, HS_ERR_INVALID, "hs_autotune_batch: empty batch");
    HS_TRY(c, hipSetDevice(c->device));
    // copies of the grooms: the candidates run on these, each from the
    // grooms' present state (restored before every candidate, so all of
    // them time the same steps)
    std::vector<hs_batch_item> tmp(items, items + n);
    std::vector<uint32_t> clones;
    std::vector<const Groom*> srcs;
    std::vector<Groom> start;  // each copy's host state at the start
    auto drop = [&] {
        (void)hipStreamSynchronize(c->stream);
        for (uint32_t id : clones) {
            Groom g;
            if (ctx_take(c, id, &g)) groom_release(g);
        }
    };
    for (uint32_t k = 0; k < n; ++k) {
        Groom* src = find(c, items[k].groom_id);
        if (!src) {
            drop();
            return fail(c, HS_ERR_NOT_FOUND, "hs_autotune_batch: unknown groom");
        }
        Groom g = *src;
        g.heap = nullptr;
        g.block = nullptr;
        hipError_t e = hipMalloc(&g.block, g.bytes);
        if (e == hipSuccess) {
            carve(g, static_cast<uint8_t*>(g.block));
            e = hipMemcpyAsync(g.block, src->block, g.bytes, hipMemcpyDeviceToDevice, c->stream);
        }
        if (e != hipSuccess) {
            if (g.block) (void)hipFree(g.block);
            drop();
            return hip_fail(c, e, "hs_autotune_batch: copy");
        }
        srcs.push_back(src);
        start.push_back(g);
        const uint32_t id = ctx_adopt(c, std::move(g));
        clones.push_back(id);
        tmp[k].groom_id = id;
    }
    auto restore = [&]() -> int {
        for (uint32_t k = 0; k < n; ++k) {
            Groom* cl = find(c, clones[k]);
            *cl = start[k];
            const hipError_t e =
                hipMemcpyAsync(cl->block, srcs[k]->block, cl->bytes, hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "hs_autotune_batch: restore");
        }
        return HS_OK;
    };
    const hs_tile saved = c->tile;
    const bool graphs = c->graphs;
    const bool prof = c->profile;
    c->profile = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = HS_OK;
    double best_t = 1e300;
    hs_tile best_tile = saved;
    int best_shape = c->shape;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = HS_ERR_HIP;
    const int saved_shape = c->shape;
    const uint32_t gcand[] = {512, 1024, 2048};
    const uint32_t scand[] = {512, 1024, 2048};
    for (int sh = 0; sh < hs::kNumShapes; ++sh)
        for (uint32_t gc : gcand)
            for (uint32_t sc : scand)
                for (uint32_t xo = 0; xo < 2 && rc == HS_OK; ++xo) {
                    if (gc > hs::kWinCells || sc > hs::kWinCells) continue;
                    c->tile = hs_tile{gc, sc, xo, hs::kShapes[sh].strands, hs::kShapes[sh].points};
                    c->shape = sh;
                    // the untimed step also takes the change of shape (its
                    // gather reads the float grid directly)
                    rc = restore();
                    if (rc == HS_OK) rc = hs_step_batch(c, tmp.data(), n);
                if (rc != HS_OK) break;
                    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = HS_ERR_HIP;
                    for (uint32_t s = 0; s < steps && rc == HS_OK; ++s) rc = hs_step_batch(c, tmp.data(), n);
                    float ms = 0.0f;
                    if (rc == HS_OK && (hipEventRecord(e1, c->stream) != hipSuccess ||
                                        hipEventSynchronize(e1) != hipSuccess ||
                                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                        rc = HS_ERR_HIP;
                    if (rc == HS_OK && ms / steps < best_t) {
                        best_t = ms / steps;
                        best_tile = c->tile;
                        best_shape = sh;
                    }
                }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    c->profile = prof;
    c->graphs = graphs;
    drop();
    if (rc != HS_OK) {
        c->tile = saved;
        c->shape = saved_shape;
        return rc;
    }
    c->tile = best_tile;
    c->shape = best_shape;
    if (best) *best = best_tile;
    if (best_ms) *best_ms = best_t;
    return HS_OK;
}

int hs_batch_stats(const hs_ctx* c, uint64_t* batch_calls, uint64_t* graph_captures)
{
    if (!c) return HS_ERR_INVALID;
    if (batch_calls) *batch_calls = c->batch_calls;
    if (graph_captures) *graph_captures = c->graph_captures;
    return HS_OK;
}

int hs_read_strands(hs_ctx* c, uint32_t id, float* aos, hs_draw_indirect* ind)
{
    if (!c) return HS_ERR_INVALID;
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_read_strands: unknown groom");
    HS_TRY(c, hipSetDevice(c->device));
    if (ind) {
        // the count compute.comp:302 accumulates: one per strand processed
        ind->vertexCount = g->S;
        ind->instanceCount = 1;
        ind->firstVertex = 0;
        ind->firstInstance = 0;
    }
    if (!aos) return HS_OK;
    const size_t bytes = (size_t)g->S * g->N * 48;
    float* dev = nullptr;
    HS_TRY(c, hipMalloc((void**)&dev, bytes));
    hipError_t e = groom_unpack(*g, dev, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(aos, dev, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(c, e, "hs_read_strands");
    return HS_OK;
}

int hs_export_strands(hs_ctx* c, uint32_t id, float* aos_dev, uint32_t* ind_dev)
{
    if (!c) return HS_ERR_INVALID;
    if (!aos_dev) return fail(c, HS_ERR_INVALID, "hs_export_strands: null buffer");
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_export_strands: unknown groom");
    HS_TRY(c, hipSetDevice(c->device));
    HS_TRY(c, groom_unpack(*g, aos_dev, c->stream));
    if (ind_dev) HS_TRY(c, hs::launch_indirect(ind_dev, g->S, c->stream));
    return HS_OK;
}

int hs_resample_strands(hs_ctx* c, uint32_t id, uint32_t M, float* out_dev)
{
    if (!c) return HS_ERR_INVALID;
    if (!out_dev) return fail(c, HS_ERR_INVALID, "hs_resample_strands: null buffer");
    if (M < 2 || M > (1u << 20)) return fail(c, HS_ERR_INVALID, "hs_resample_strands: samples must be in [2, 2^20]");
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_resample_strands: unknown groom");
    HS_TRY(c, hipSetDevice(c->device));
    HS_TRY(c, hs::launch_resample(g->pos, g->perm, g->S, g->N, g->L.Sp, M, out_dev, c->stream));
    return HS_OK;
}

int hs_write_strands(hs_ctx* c, uint32_t id, const float* aos)
{
    if (!c) return HS_ERR_INVALID;
    if (!aos) return fail(c, HS_ERR_INVALID, "hs_write_strands: null buffer");
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_write_strands: unknown groom");
    HS_TRY(c, hipSetDevice(c->device));
    const size_t bytes = (size_t)g->S * g->N * 48;
    float* dev = nullptr;
    HS_TRY(c, hipMalloc((void**)&dev, bytes));
    hipError_t e = hipMemcpyAsync(dev, aos, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hs::launch_pack(dev, g->perm, g->S, g->N, g->L.Sp, g->pos, g->vel, g->corr, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(c, e, "hs_write_strands");
    g->pending = false;
    return HS_OK;
}

int hs_read_grid(hs_ctx* c, uint32_t id, hs_grid_cell* cells)
{
    if (!c) return HS_ERR_INVALID;
    if (!cells) return fail(c, HS_ERR_INVALID, "hs_read_grid: null buffer");
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_read_grid: unknown groom");
    if (g->G == 0) return fail(c, HS_ERR_STATE, "hs_read_grid: groom has no grid");
    HS_TRY(c, hipSetDevice(c->device));
    const size_t bytes = grid_cells(g->G) * 16;
    if (!g->pending) {
        std::memset(cells, 0, bytes);
        return HS_OK;
    }
    HS_TRY(c, hipMemcpyAsync(cells, g->gw.grid_i, bytes, hipMemcpyDeviceToHost, c->stream));
    HS_TRY(c, hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_grid_work(hs_ctx* c, uint32_t id, hs_grid_work_stats* out)
{
    if (!c) return HS_ERR_INVALID;
    if (!out) return fail(c, HS_ERR_INVALID, "hs_grid_work: null output");
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_grid_work: unknown groom");
    std::memset(out, 0, sizeof(*out));
    if (g->G == 0 || !g->pending) return HS_OK;
    HS_TRY(c, hipSetDevice(c->device));
    uint32_t ctr[4] = {0, 0, 0, 0};
    // the last splat wrote the half before the current parity
    HS_TRY(c, hipMemcpyAsync(ctr, g->gw.counters + (g->parity ^ 1u) * 4, sizeof(ctr), hipMemcpyDeviceToHost,
                             c->stream));
    HS_TRY(c, hipStreamSynchronize(c->stream));
    out->pool_cells = ctr[0];
    out->pool_capacity = g->gw.pool_cap - g->gw.nws * hs::kSlotCells;
    out->boxes = ctr[1];
    out->box_capacity = g->gw.desc_cap - g->gw.nws;
    out->fallback = ctr[2];
    out->deferred = ctr[3];
    return HS_OK;
}

int hs_grid_near_limit(hs_ctx* c, uint32_t id, uint64_t* cells)
{
    if (!c) return HS_ERR_INVALID;
    if (!cells) return fail(c, HS_ERR_INVALID, "hs_grid_near_limit: null output");
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_grid_near_limit: unknown groom");
    *cells = 0;
    if (g->G == 0) return HS_OK;
    HS_TRY(c, hipSetDevice(c->device));
    unsigned long long v = 0;
    HS_TRY(c, hipMemcpyAsync(&v, g->gw.near_limit, 8, hipMemcpyDeviceToHost, c->stream));
    HS_TRY(c, hipStreamSynchronize(c->stream));
    *cells = v;
    return HS_OK;
}

int hs_profile_enable(hs_ctx* c, int on)
{
    if (!c) return HS_ERR_INVALID;
    int rc = drain_profile(c);
    if (rc != HS_OK) return rc;
    c->profile = on != 0;
    c->prof_launches = 0;
    c->prof_step_ms = 0.0;
    c->prof_reduce_ms = 0.0;
    c->prof_mesh_ms = 0.0;
    return HS_OK;
}

int hs_profile_read(hs_ctx* c, uint64_t* launches, double* total_ms)
{
    if (!c || !launches || !total_ms) return HS_ERR_INVALID;
    int rc = drain_profile(c);
    if (rc != HS_OK) return rc;
    *launches = c->prof_launches;
    *total_ms = c->prof_step_ms + c->prof_reduce_ms + c->prof_mesh_ms;
    return HS_OK;
}

int hs_profile_breakdown(hs_ctx* c, double* step_kernel_ms, double* reduce_kernel_ms)
{
    if (!c || !step_kernel_ms || !reduce_kernel_ms) return HS_ERR_INVALID;
    int rc = drain_profile(c);
    if (rc != HS_OK) return rc;
    *step_kernel_ms = c->prof_step_ms;
    *reduce_kernel_ms = c->prof_reduce_ms;
    return HS_OK;
}

int hs_profile_mesh(hs_ctx* c, double* mesh_kernel_ms)
{
    if (!c || !mesh_kernel_ms) return HS_ERR_INVALID;
    int rc = drain_profile(c);
    if (rc != HS_OK) return rc;
    *mesh_kernel_ms = c->prof_mesh_ms;
    return HS_OK;
}

}  // extern "C"
