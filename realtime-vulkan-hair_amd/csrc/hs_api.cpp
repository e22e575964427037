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
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/hairsim.h"
#include "hs_kernels.h"

namespace {

uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

struct Groom {
    uint32_t S = 0, N = 0, Sp = 0, G = 0, flags = 0;
    float origin[3] = {0, 0, 0};
    float h = 1.0f;
    float* pos = nullptr;
    float* vel = nullptr;
    float* corr = nullptr;
    float* seg = nullptr;
    float* noise = nullptr;
    float noise_dir[3] = {0, 0, 0};
    int4* grid[3] = {nullptr, nullptr, nullptr};
    bool dirty[3] = {false, false, false};
    // Grid friction of the last step, not yet applied to the stored velocities.
    int pending = -1;
    float pend_mu = 0.0f, pend_omu = 1.0f;
    uint64_t bytes = 0;
};

}  // namespace

struct hs_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<uint32_t, Groom> grooms;
    uint32_t next_id = 1;
    uint64_t bytes = 0;
    // kernel timing (pooled events around each step launch on the context stream)
    bool profile = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t events_used = 0;
    uint64_t prof_launches = 0;
    double prof_ms = 0.0;
};

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

#define HS_TRY(ctx, expr)                                     \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

void free_groom(hs_ctx* c, Groom& g)
{
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(g.pos);
    (void)hipFree(g.vel);
    (void)hipFree(g.corr);
    (void)hipFree(g.seg);
    (void)hipFree(g.noise);
    for (int k = 0; k < 3; ++k) (void)hipFree(g.grid[k]);
    c->bytes -= g.bytes;
    g = Groom();
}

Groom* find(hs_ctx* c, uint32_t id)
{
    auto it = c->grooms.find(id);
    return it == c->grooms.end() ? nullptr : &it->second;
}

uint64_t grid_cells(uint32_t G) { return (uint64_t)G * G * G; }

int drain_profile(hs_ctx* c)
{
    for (size_t k = 0; k < c->events_used; ++k) {
        auto& ev = c->events[k];
        HS_TRY(c, hipEventSynchronize(ev.second));
        float ms = 0.0f;
        HS_TRY(c, hipEventElapsedTime(&ms, ev.first, ev.second));
        c->prof_ms += ms;
        c->prof_launches += 1;
    }
    c->events_used = 0;
    return HS_OK;
}

// Next pair of timing events from the pool (created on first use).
int next_events(hs_ctx* c, hipEvent_t* e0, hipEvent_t* e1)
{
    if (c->events_used == c->events.size()) {
        if (c->events.size() >= 4096) {
            int rc = drain_profile(c);
            if (rc != HS_OK) return rc;
        } else {
            hipEvent_t a = nullptr, b = nullptr;
            HS_TRY(c, hipEventCreate(&a));
            HS_TRY(c, hipEventCreate(&b));
            c->events.emplace_back(a, b);
        }
    }
    *e0 = c->events[c->events_used].first;
    *e1 = c->events[c->events_used].second;
    c->events_used += 1;
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
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
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
    for (auto& ev : c->events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
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
    const uint64_t Sp = round_up(S, 64);
    uint64_t b = 2ull * 3 * N * Sp * 4 + 2ull * Sp * 4;
    if (flags & HS_GROOM_KEEP_CORRECTIONS) b += 3ull * N * Sp * 4;
    b += 3ull * grid_cells(G) * 16;
    return b;
}

int hs_groom_create(hs_ctx* c, const hs_groom_desc* d, const float* aos, uint32_t* id)
{
    if (!c) return HS_ERR_INVALID;
    if (!d || !aos || !id) return fail(c, HS_ERR_INVALID, "hs_groom_create: null argument");
    if (d->num_strands == 0) return fail(c, HS_ERR_INVALID, "hs_groom_create: num_strands must be >= 1");
    if (d->num_points < 2 || d->num_points > HS_MAX_POINTS)
        return fail(c, HS_ERR_INVALID, "hs_groom_create: num_points must be in [2, 256]");
    if (d->num_strands > (1u << 26)) return fail(c, HS_ERR_INVALID, "hs_groom_create: too many strands");
    if (d->grid_dim > 1024) return fail(c, HS_ERR_INVALID, "hs_groom_create: grid_dim > 1024");
    if (d->grid_dim > 0 && !(d->grid_h > 0.0f)) return fail(c, HS_ERR_INVALID, "hs_groom_create: grid_h must be > 0");
    if (!d->seg_length && !(d->strand_length > 0.0f))
        return fail(c, HS_ERR_INVALID, "hs_groom_create: strand_length must be > 0");
    HS_TRY(c, hipSetDevice(c->device));

    Groom g;
    g.S = d->num_strands;
    g.N = d->num_points;
    g.Sp = round_up(g.S, 64);
    g.G = d->grid_dim;
    g.flags = d->flags;
    g.h = d->grid_h;
    for (int k = 0; k < 3; ++k) {
        g.origin[k] = d->grid_origin[k];
        g.noise_dir[k] = d->noise_dir[k];
    }
    const size_t planes = 3ull * g.N * g.Sp;
    const size_t aos_bytes = (size_t)g.S * g.N * 48;
    float* staging = nullptr;
    auto cleanup = [&](int code, const std::string& msg) {
        (void)hipFree(staging);
        (void)hipFree(g.pos);
        (void)hipFree(g.vel);
        (void)hipFree(g.corr);
        (void)hipFree(g.seg);
        (void)hipFree(g.noise);
        for (int k = 0; k < 3; ++k) (void)hipFree(g.grid[k]);
        return fail(c, code, msg);
    };
#define HS_ALLOC(ptr, bytes)                                                          \
    do {                                                                              \
        hipError_t e_ = hipMalloc((void**)&(ptr), (bytes));                           \
        if (e_ != hipSuccess)                                                         \
            return cleanup(e_ == hipErrorOutOfMemory ? HS_ERR_OUT_OF_MEMORY : HS_ERR_HIP, \
                           std::string("hipMalloc: ") + hipGetErrorString(e_));       \
    } while (0)
#define HS_CHECK(expr)                                                                \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) return cleanup(HS_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

    HS_ALLOC(g.pos, planes * 4);
    HS_ALLOC(g.vel, planes * 4);
    if (g.flags & HS_GROOM_KEEP_CORRECTIONS) HS_ALLOC(g.corr, planes * 4);
    HS_ALLOC(g.seg, (size_t)g.Sp * 4);
    if (d->wind_noise) HS_ALLOC(g.noise, (size_t)g.Sp * 4);
    for (int k = 0; k < 3 && g.G; ++k) HS_ALLOC(g.grid[k], grid_cells(g.G) * 16);
    HS_ALLOC(staging, aos_bytes);

    // Padding lanes (s >= S) are never stepped; keep them zero.
    HS_CHECK(hipMemsetAsync(g.pos, 0, planes * 4, c->stream));
    HS_CHECK(hipMemsetAsync(g.vel, 0, planes * 4, c->stream));
    if (g.corr) HS_CHECK(hipMemsetAsync(g.corr, 0, planes * 4, c->stream));
    for (int k = 0; k < 3 && g.G; ++k) HS_CHECK(hipMemsetAsync(g.grid[k], 0, grid_cells(g.G) * 16, c->stream));

    std::vector<float> seg(g.Sp, 0.0f);
    if (d->seg_length) {
        for (uint32_t s = 0; s < g.S; ++s) seg[s] = d->seg_length[s];
    } else {
        // radius = strandLength / (NUM_CURVE_POINTS - 1.0)  (compute.comp:139-140)
        const float r = d->strand_length / ((float)g.N - 1.0f);
        for (uint32_t s = 0; s < g.S; ++s) seg[s] = r;
    }
    HS_CHECK(hipMemcpyAsync(g.seg, seg.data(), (size_t)g.Sp * 4, hipMemcpyHostToDevice, c->stream));
    if (d->wind_noise) {
        std::vector<float> nz(g.Sp, 0.0f);
        std::copy(d->wind_noise, d->wind_noise + g.S, nz.begin());
        HS_CHECK(hipMemcpyAsync(g.noise, nz.data(), (size_t)g.Sp * 4, hipMemcpyHostToDevice, c->stream));
        HS_CHECK(hipStreamSynchronize(c->stream));
    }
    HS_CHECK(hipMemcpyAsync(staging, aos, aos_bytes, hipMemcpyHostToDevice, c->stream));
    HS_CHECK(hs::launch_pack(staging, g.S, g.N, g.Sp, g.pos, g.vel, g.corr, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    HS_CHECK(hipFree(staging));
    staging = nullptr;
#undef HS_ALLOC
#undef HS_CHECK

    g.bytes = hs_groom_bytes(g.S, g.N, g.G, g.flags) - (g.noise ? 0 : (uint64_t)g.Sp * 4);
    c->bytes += g.bytes;
    const uint32_t gid = c->next_id++;
    c->grooms[gid] = g;
    *id = gid;
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
    if (!t || !p) return fail(c, HS_ERR_INVALID, "hs_step: null time or params");
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_step: unknown groom");
    if (ncol > HS_MAX_COLLIDERS) return fail(c, HS_ERR_INVALID, "hs_step: more than 8 colliders");
    if (ncol > 0 && !cols) return fail(c, HS_ERR_INVALID, "hs_step: null colliders");
    if (p->num_colliders > ncol) return fail(c, HS_ERR_INVALID, "hs_step: params.num_colliders > colliders given");
    if (!(t->deltaTime > 0.0f) || !std::isfinite(t->deltaTime))
        return fail(c, HS_ERR_INVALID, "hs_step: deltaTime must be finite and > 0");
    if (p->mesh_id != 0) return fail(c, HS_ERR_NOT_FOUND, "hs_step: unknown mesh");
    HS_TRY(c, hipSetDevice(c->device));

    hs::StepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.pos = g->pos;
    a.vel = g->vel;
    a.corr = g->corr;
    a.seg = g->seg;
    a.noise = g->noise;
    a.S = g->S;
    a.N = g->N;
    a.Sp = g->Sp;
    a.G = g->G;
    a.ox = g->origin[0];
    a.oy = g->origin[1];
    a.oz = g->origin[2];
    a.inv_h = g->G ? 1.0f / g->h : 0.0f;
    a.dt = t->deltaTime;
    a.inv_dt = 1.0f / a.dt;
    a.dt2 = a.dt * a.dt;
    a.D = p->damping;
    a.k = p->spring_k;
    a.vmax = p->vmax;
    a.sphere_r = p->sphere_radius;
    a.fx = p->gravity[0] + p->wind[0];
    a.fy = p->gravity[1] + p->wind[1];
    a.fz = p->gravity[2] + p->wind[2];
    a.ndx = g->noise_dir[0];
    a.ndy = g->noise_dir[1];
    a.ndz = g->noise_dir[2];
    a.ncol = p->num_colliders;
    for (uint32_t j = 0; j < a.ncol; ++j) {
        std::memcpy(a.col[j].m, cols[j].transform, 64);
        std::memcpy(a.col[j].m + 16, cols[j].inv, 64);
        std::memcpy(a.col[j].m + 32, cols[j].invTrans, 64);
    }

    const bool gather = g->pending >= 0;
    if (gather) {
        a.gprev = g->grid[g->pending];
        a.mu_prev = g->pend_mu;
        a.omu_prev = g->pend_omu;
    }
    const bool splat = p->hair_hair && g->G > 0;
    int next = -1;
    if (splat) {
        // the splat target is any buffer that is not pending; it must be zero
        next = (g->pending + 1 + 3) % 3;
        if (g->dirty[next]) {
            HS_TRY(c, hipMemsetAsync(g->grid[next], 0, grid_cells(g->G) * 16, c->stream));
            g->dirty[next] = false;
        }
        // a buffer that is neither pending nor the target is cleared in-kernel
        // for the following step
        int clr = -1;
        for (int k = 0; k < 3; ++k)
            if (k != next && k != g->pending) clr = k;
        if (clr >= 0 && g->dirty[clr]) {
            a.gclear = g->grid[clr];
            a.clear_cells = (uint32_t)grid_cells(g->G);
            g->dirty[clr] = false;
        }
        a.gnext = g->grid[next];
        g->dirty[next] = true;
    }

    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->profile) {
        int rc = next_events(c, &e0, &e1);
        if (rc != HS_OK) return rc;
        HS_TRY(c, hipEventRecord(e0, c->stream));
    }
    HS_TRY(c, hs::launch_step(a, gather, splat, c->stream));
    if (c->profile) HS_TRY(c, hipEventRecord(e1, c->stream));

    if (splat) {
        g->pending = next;
        g->pend_mu = p->friction;
        g->pend_omu = 1.0f - p->friction;
    } else {
        g->pending = -1;
    }
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
    hs::UnpackGrid ug;
    const hs::UnpackGrid* pug = nullptr;
    if (g->pending >= 0) {
        ug.grid = g->grid[g->pending];
        ug.G = g->G;
        ug.ox = g->origin[0];
        ug.oy = g->origin[1];
        ug.oz = g->origin[2];
        ug.inv_h = 1.0f / g->h;
        ug.mu = g->pend_mu;
        ug.omu = g->pend_omu;
        pug = &ug;
    }
    hipError_t e = hs::launch_unpack(dev, g->S, g->N, g->Sp, g->pos, g->vel, g->corr, pug, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(aos, dev, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(c, e, "hs_read_strands");
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
    if (e == hipSuccess) e = hs::launch_pack(dev, g->S, g->N, g->Sp, g->pos, g->vel, g->corr, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(c, e, "hs_write_strands");
    g->pending = -1;
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
    if (g->pending < 0) {
        std::memset(cells, 0, bytes);
        return HS_OK;
    }
    HS_TRY(c, hipMemcpyAsync(cells, g->grid[g->pending], bytes, hipMemcpyDeviceToHost, c->stream));
    HS_TRY(c, hipStreamSynchronize(c->stream));
    return HS_OK;
}

int hs_profile_enable(hs_ctx* c, int on)
{
    if (!c) return HS_ERR_INVALID;
    int rc = drain_profile(c);
    if (rc != HS_OK) return rc;
    c->profile = on != 0;
    c->prof_launches = 0;
    c->prof_ms = 0.0;
    return HS_OK;
}

int hs_profile_read(hs_ctx* c, uint64_t* launches, double* total_ms)
{
    if (!c || !launches || !total_ms) return HS_ERR_INVALID;
    int rc = drain_profile(c);
    if (rc != HS_OK) return rc;
    *launches = c->prof_launches;
    *total_ms = c->prof_ms;
    return HS_OK;
}

}  // extern "C"
