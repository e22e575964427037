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

namespace {

uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }
uint64_t grid_cells(uint32_t G) { return (uint64_t)G * G * G; }

// Shape-derived sizes of one groom's device state.
struct Layout {
    uint32_t Sp, nwg, nseg;
    uint64_t pool_cap, desc_cap;
    uint64_t state_bytes, grid_bytes;
};

Layout layout_of(uint32_t S, uint32_t N, uint32_t G, uint32_t flags, bool noise)
{
    Layout l;
    l.Sp = (uint32_t)round_up(S, hs::kW);
    l.nwg = l.Sp / hs::kW;
    l.nseg = (N - 1 + hs::kL - 1) / hs::kL;
    const uint64_t planes = 3ull * N * l.Sp * 4;
    l.state_bytes = 2 * planes + (uint64_t)l.Sp * 4 /*seg*/ + (uint64_t)l.Sp * 4 /*perm*/ +
                    sizeof(hs::Collider) * hs::kMaxColliders;
    if (noise) l.state_bytes += (uint64_t)l.Sp * 4;
    if (flags & HS_GROOM_KEEP_CORRECTIONS) l.state_bytes += planes;
    l.pool_cap = 0;
    l.desc_cap = 0;
    l.grid_bytes = 0;
    if (G) {
        // partial boxes: room for half a cell per point per step, at least 64k cells
        l.pool_cap = std::max<uint64_t>(65536, (uint64_t)S * (N - 1) / 2);
        l.desc_cap = 4ull * l.nwg * l.nseg + 4096;
        l.grid_bytes = 3 * grid_cells(G) * 16                 // grid_i, grid_f, overflow
                       + 2 * l.pool_cap * 16                  // pool halves
                       + 2 * l.desc_cap * hs::kDescInts * 4   // descriptor halves
                       + 8 * 4                                // counters
                       + (uint64_t)l.nwg * l.nseg * 6 * 4;    // workgroup-segment boxes
    }
    return l;
}

struct Groom {
    uint32_t S = 0, N = 0, G = 0, flags = 0;
    Layout L{};
    float origin[3] = {0, 0, 0};
    float h = 1.0f;
    float noise_dir[3] = {0, 0, 0};
    float* pos = nullptr;
    float* vel = nullptr;
    float* corr = nullptr;
    float* seg = nullptr;
    float* noise = nullptr;
    uint32_t* perm = nullptr;   // physical slot -> caller's strand index
    hs::Collider* cols = nullptr;  // device copy of the step's colliders (ellipsoids)
    hs::GridWork gw{};
    std::vector<void*> allocs;  // every device allocation of this groom
    // Grid friction of the last step, not yet applied to the stored velocities.
    bool pending = false;
    float pend_mu = 0.0f, pend_omu = 1.0f;
    uint32_t parity = 0;        // pool/counter half the next splat writes
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
    // kernel timing: pooled events (step start, step end, reduce end) per step
    bool profile = false;
    struct Ev {
        hipEvent_t e[3];
        bool reduce;
    };
    std::vector<Ev> events;
    size_t events_used = 0;
    uint64_t prof_launches = 0;
    double prof_step_ms = 0.0, prof_reduce_ms = 0.0;
    // device staging of hs_step_batch arguments
    hs::StepArgs* batch_args = nullptr;
    uint32_t* batch_idx = nullptr;
    size_t batch_args_cap = 0, batch_idx_cap = 0;
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

#define HS_TRY(ctx, expr)                                      \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

void free_groom(hs_ctx* c, Groom& g)
{
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (void* p : g.allocs) (void)hipFree(p);
    c->bytes -= g.bytes;
    g = Groom();
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
        if (ev.reduce) {
            HS_TRY(c, hipEventElapsedTime(&ms, ev.e[1], ev.e[2]));
            c->prof_reduce_ms += ms;
        }
        c->prof_launches += 1;
    }
    c->events_used = 0;
    return HS_OK;
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
            for (int k = 0; k < 3; ++k) HS_TRY(c, hipEventCreate(&ev.e[k]));
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
// exactly kW strands (one workgroup each; only the last leaf may be short).
// Every workgroup's strands start close together, so the voxels they touch
// form a compact box; unlike a space-filling curve there is no jump between
// distant patches inside one workgroup.
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
        if (n <= hs::kW) continue;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = r.b; i < r.e; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], root[3ull * idx[i] + k]);
                hi[k] = std::max(hi[k], root[3ull * idx[i] + k]);
            }
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
        // split at the multiple of kW nearest the median: the left part is
        // always whole workgroups, the remainder travels right
        uint32_t mid = (uint32_t)(((n / 2) + hs::kW / 2) / hs::kW) * hs::kW;
        mid = std::min(std::max(mid, (uint32_t)hs::kW), n - 1);
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
int build_args(hs_ctx* c, Groom& g, const hs_time* t, const hs_collider* cols, uint32_t ncol,
               const hs_step_params* p, const char* who, hs::StepArgs* out, bool* gather, bool* splat)
{
    const std::string w(who);
    if (!t || !p) return fail(c, HS_ERR_INVALID, w + ": null time or params");
    if (ncol > HS_MAX_COLLIDERS) return fail(c, HS_ERR_INVALID, w + ": more than 8 colliders");
    if (ncol > 0 && !cols) return fail(c, HS_ERR_INVALID, w + ": null colliders");
    if (p->num_colliders > ncol) return fail(c, HS_ERR_INVALID, w + ": params.num_colliders > colliders given");
    if (!(t->deltaTime > 0.0f) || !std::isfinite(t->deltaTime))
        return fail(c, HS_ERR_INVALID, w + ": deltaTime must be finite and > 0");
    if (p->mesh_id != 0) return fail(c, HS_ERR_NOT_FOUND, w + ": unknown mesh");
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
    a.nwg = g.L.nwg;
    a.nseg = g.L.nseg;
    a.gw = g.gw;
    a.G = g.G;
    a.parity = g.parity;
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
    a.ncol = p->num_colliders;
    a.cols = g.cols;
    if (a.ncol > 0) {
        a.sphere_c[0] = cols[0].transform[12];
        a.sphere_c[1] = cols[0].transform[13];
        a.sphere_c[2] = cols[0].transform[14];
    }
    if (a.ncol > 1) {
        // ellipsoids are read from device memory; stream order keeps the
        // previous step's kernel ahead of this copy
        hs::Collider host[hs::kMaxColliders];
        for (uint32_t j = 0; j < a.ncol; ++j) {
            std::memcpy(host[j].m, cols[j].transform, 64);
            std::memcpy(host[j].m + 16, cols[j].inv, 64);
            std::memcpy(host[j].m + 32, cols[j].invTrans, 64);
        }
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

// Host-side state after a launched step: pending friction and pool parity.
void advance_state(Groom& g, bool splat, float friction)
{
    if (splat) {
        g.parity ^= 1u;
        g.pending = true;
        g.pend_mu = friction;
        g.pend_omu = 1.0f - friction;
    } else {
        g.pending = false;
    }
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
    for (auto& ev : c->events)
        for (int k = 0; k < 3; ++k) (void)hipEventDestroy(ev.e[k]);
    (void)hipFree(c->batch_args);
    (void)hipFree(c->batch_idx);
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
    const Layout l = layout_of(S, N, G, flags, true);
    return l.state_bytes + l.grid_bytes;
}

int hs_groom_create(hs_ctx* c, const hs_groom_desc* d, const float* aos, uint32_t* id)
{
    if (!c) return HS_ERR_INVALID;
    if (!d || !aos || !id) return fail(c, HS_ERR_INVALID, "hs_groom_create: null argument");
    if (d->num_strands == 0) return fail(c, HS_ERR_INVALID, "hs_groom_create: num_strands must be >= 1");
    if (d->num_points < 2 || d->num_points > HS_MAX_POINTS)
        return fail(c, HS_ERR_INVALID, "hs_groom_create: num_points must be in [2, 256]");
    if (d->num_strands > (1u << 26)) return fail(c, HS_ERR_INVALID, "hs_groom_create: too many strands");
    // kernels address the [3][N][Sp] state with 32-bit byte offsets: 12*N*Sp < 2^32
    if ((uint64_t)d->num_points * round_up(d->num_strands, hs::kW) * 12 >= (1ull << 32))
        return fail(c, HS_ERR_INVALID, "hs_groom_create: more than 357M points in one groom");
    if (d->grid_dim > 1024) return fail(c, HS_ERR_INVALID, "hs_groom_create: grid_dim > 1024");
    if (d->grid_dim > 0 && !(d->grid_h > 0.0f && std::isfinite(d->grid_h)))
        return fail(c, HS_ERR_INVALID, "hs_groom_create: grid_h must be finite and > 0");
    if (!d->seg_length && !(d->strand_length > 0.0f))
        return fail(c, HS_ERR_INVALID, "hs_groom_create: strand_length must be > 0");
    HS_TRY(c, hipSetDevice(c->device));

    Groom g;
    g.S = d->num_strands;
    g.N = d->num_points;
    g.G = d->grid_dim;
    g.flags = d->flags;
    g.h = d->grid_h;
    for (int k = 0; k < 3; ++k) {
        g.origin[k] = d->grid_origin[k];
        g.noise_dir[k] = d->noise_dir[k];
    }
    g.L = layout_of(g.S, g.N, g.G, g.flags, d->wind_noise != nullptr);
    const Layout& L = g.L;
    const size_t planes = 3ull * g.N * L.Sp;
    const size_t aos_bytes = (size_t)g.S * g.N * 48;
    float* staging = nullptr;
    auto cleanup = [&](int code, const std::string& msg) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(staging);
        for (void* p : g.allocs) (void)hipFree(p);
        return fail(c, code, msg);
    };
#define HS_ALLOC(ptr, nbytes)                                                                     \
    do {                                                                                          \
        void* p_ = nullptr;                                                                       \
        hipError_t e_ = hipMalloc(&p_, (nbytes));                                                 \
        if (e_ != hipSuccess)                                                                     \
            return cleanup(e_ == hipErrorOutOfMemory ? HS_ERR_OUT_OF_MEMORY : HS_ERR_HIP,          \
                           std::string("hipMalloc: ") + hipGetErrorString(e_));                   \
        g.allocs.push_back(p_);                                                                   \
        (ptr) = reinterpret_cast<decltype(ptr)>(p_);                                              \
        e_ = hipMemsetAsync(p_, 0, (nbytes), c->stream);                                          \
        if (e_ != hipSuccess) return cleanup(HS_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e_)); \
    } while (0)
#define HS_CHECK(expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return cleanup(HS_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

    HS_ALLOC(g.pos, planes * 4);
    HS_ALLOC(g.vel, planes * 4);
    if (g.flags & HS_GROOM_KEEP_CORRECTIONS) HS_ALLOC(g.corr, planes * 4);
    HS_ALLOC(g.seg, (size_t)L.Sp * 4);
    HS_ALLOC(g.perm, (size_t)L.Sp * 4);
    HS_ALLOC(g.cols, sizeof(hs::Collider) * hs::kMaxColliders);
    if (d->wind_noise) HS_ALLOC(g.noise, (size_t)L.Sp * 4);
    if (g.G) {
        const size_t cells = grid_cells(g.G);
        HS_ALLOC(g.gw.grid_i, cells * 16);
        HS_ALLOC(g.gw.grid_f, cells * 16);
        HS_ALLOC(g.gw.overflow, cells * 16);
        HS_ALLOC(g.gw.pool, 2 * L.pool_cap * 16);
        HS_ALLOC(g.gw.desc, 2 * L.desc_cap * hs::kDescInts * 4);
        HS_ALLOC(g.gw.counters, 8 * 4);
        HS_ALLOC(g.gw.gbox, (size_t)L.nwg * L.nseg * 6 * 4);
        g.gw.pool_cap = (uint32_t)L.pool_cap;
        g.gw.desc_cap = (uint32_t)L.desc_cap;
    }
    HS_CHECK(hipMalloc((void**)&staging, aos_bytes));

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
    HS_CHECK(hipMemcpyAsync(g.seg, seg.data(), (size_t)L.Sp * 4, hipMemcpyHostToDevice, c->stream));
    HS_CHECK(hipMemcpyAsync(g.perm, perm.data(), (size_t)g.S * 4, hipMemcpyHostToDevice, c->stream));
    if (d->wind_noise) HS_CHECK(hipMemcpyAsync(g.noise, nz.data(), (size_t)L.Sp * 4, hipMemcpyHostToDevice, c->stream));
    HS_CHECK(hipMemcpyAsync(staging, aos, aos_bytes, hipMemcpyHostToDevice, c->stream));
    HS_CHECK(hs::launch_pack(staging, g.perm, g.S, g.N, L.Sp, g.pos, g.vel, g.corr, c->stream));
    HS_CHECK(hipStreamSynchronize(c->stream));
    HS_CHECK(hipFree(staging));
    staging = nullptr;
#undef HS_ALLOC
#undef HS_CHECK

    g.bytes = L.state_bytes + L.grid_bytes;
    c->bytes += g.bytes;
    const uint32_t gid = c->next_id++;
    c->grooms[gid] = std::move(g);
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
    Groom* g = find(c, id);
    if (!g) return fail(c, HS_ERR_NOT_FOUND, "hs_step: unknown groom");
    HS_TRY(c, hipSetDevice(c->device));
    hs::StepArgs a;
    bool gather = false, splat = false;
    int rc = build_args(c, *g, t, cols, ncol, p, "hs_step", &a, &gather, &splat);
    if (rc != HS_OK) return rc;

    hs_ctx::Ev* ev = nullptr;
    if (c->profile) {
        rc = next_events(c, &ev);
        if (rc != HS_OK) return rc;
        ev->reduce = splat;
        HS_TRY(c, hipEventRecord(ev->e[0], c->stream));
    }
    HS_TRY(c, hs::launch_step(a, gather, splat, c->stream));
    if (ev) HS_TRY(c, hipEventRecord(ev->e[1], c->stream));
    if (splat) {
        HS_TRY(c, hs::launch_reduce(g->gw, g->G, g->parity, c->stream));
        if (ev) HS_TRY(c, hipEventRecord(ev->e[2], c->stream));
    }
    advance_state(*g, splat, p->friction);
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
        for (uint32_t j = 0; j < k; ++j)
            if (gs[j] == gs[k]) return fail(c, HS_ERR_INVALID, "hs_step_batch: groom listed twice");
    }
    std::vector<hs::StepArgs> args(n);
    std::vector<uint8_t> variant(n);
    for (uint32_t k = 0; k < n; ++k) {
        bool gather = false, splat = false;
        int rc = build_args(c, *gs[k], &items[k].time, items[k].colliders, items[k].num_colliders, items[k].params,
                            "hs_step_batch", &args[k], &gather, &splat);
        if (rc != HS_OK) return rc;
        variant[k] = (uint8_t)((gather ? 1 : 0) | (splat ? 2 : 0) | (args[k].ncol > 1 ? 4 : 0));
    }
    // one launch pair per kernel variant present (usually exactly one)
    hs_ctx::Ev* ev = nullptr;
    bool any_splat = false;
    for (uint32_t k = 0; k < n; ++k) any_splat |= (variant[k] & 2) != 0;
    if (c->profile) {
        int rc = next_events(c, &ev);
        if (rc != HS_OK) return rc;
        ev->reduce = any_splat;
        HS_TRY(c, hipEventRecord(ev->e[0], c->stream));
    }
    std::vector<std::vector<uint32_t>> groups(8);
    for (uint32_t k = 0; k < n; ++k) groups[variant[k]].push_back(k);
    // host staging: args of every group, then the prefix arrays
    std::vector<hs::StepArgs> hargs;
    std::vector<uint32_t> hidx;
    struct Launch {
        size_t args_off, wg_off, br_off;
        uint32_t ng, total_wg, total_br;
        uint8_t v;
    };
    std::vector<Launch> launches;
    for (uint8_t v = 0; v < 8; ++v) {
        if (groups[v].empty()) continue;
        Launch L{hargs.size(), 0, 0, (uint32_t)groups[v].size(), 0, 0, v};
        std::vector<uint32_t> wg{0}, br{0};
        for (uint32_t k : groups[v]) {
            hargs.push_back(args[k]);
            wg.push_back(wg.back() + args[k].nwg);
            const uint32_t nb = args[k].G ? (args[k].G + hs::kBrick - 1) / hs::kBrick : 0;
            br.push_back(br.back() + ((v & 2) ? nb * nb * nb : 0));
        }
        L.total_wg = wg.back();
        L.total_br = br.back();
        L.wg_off = hidx.size();
        hidx.insert(hidx.end(), wg.begin(), wg.end());
        L.br_off = hidx.size();
        hidx.insert(hidx.end(), br.begin(), br.end());
        launches.push_back(L);
    }
    const size_t need_args = hargs.size() * sizeof(hs::StepArgs);
    const size_t need_idx = hidx.size() * sizeof(uint32_t);
    if (need_args > c->batch_args_cap || need_idx > c->batch_idx_cap) {
        HS_TRY(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->batch_args);
        (void)hipFree(c->batch_idx);
        c->batch_args = nullptr;
        c->batch_idx = nullptr;
        c->batch_args_cap = c->batch_idx_cap = 0;
        HS_TRY(c, hipMalloc((void**)&c->batch_args, std::max<size_t>(need_args, 4096) * 2));
        HS_TRY(c, hipMalloc((void**)&c->batch_idx, std::max<size_t>(need_idx, 4096) * 2));
        c->batch_args_cap = std::max<size_t>(need_args, 4096) * 2;
        c->batch_idx_cap = std::max<size_t>(need_idx, 4096) * 2;
    }
    // pageable sources: the copies are staged before returning, and stream
    // order keeps the previous batch's kernels ahead of them
    HS_TRY(c, hipMemcpyAsync(c->batch_args, hargs.data(), need_args, hipMemcpyHostToDevice, c->stream));
    HS_TRY(c, hipMemcpyAsync(c->batch_idx, hidx.data(), need_idx, hipMemcpyHostToDevice, c->stream));
    for (const Launch& L : launches) {
        const hs::StepArgs* ad = c->batch_args + L.args_off;
        HS_TRY(c, hs::launch_step_batch(ad, c->batch_idx + L.wg_off, L.ng, L.total_wg, (L.v & 1) != 0,
                                        (L.v & 2) != 0, (L.v & 4) != 0, c->stream));
    }
    if (ev) HS_TRY(c, hipEventRecord(ev->e[1], c->stream));
    for (const Launch& L : launches)
        if (L.v & 2)
            HS_TRY(c, hs::launch_reduce_batch(c->batch_args + L.args_off, c->batch_idx + L.br_off, L.ng,
                                              L.total_br, c->stream));
    if (ev && any_splat) HS_TRY(c, hipEventRecord(ev->e[2], c->stream));
    for (uint32_t k = 0; k < n; ++k) advance_state(*gs[k], (variant[k] & 2) != 0, items[k].params->friction);
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
    if (g->pending) {
        ug.grid = g->gw.grid_i;
        ug.G = g->G;
        ug.ox = g->origin[0];
        ug.oy = g->origin[1];
        ug.oz = g->origin[2];
        ug.inv_h = 1.0f / g->h;
        ug.mu = g->pend_mu;
        ug.omu = g->pend_omu;
        pug = &ug;
    }
    hipError_t e = hs::launch_unpack(dev, g->perm, g->S, g->N, g->L.Sp, g->pos, g->vel, g->corr, pug, c->stream);
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

int hs_profile_enable(hs_ctx* c, int on)
{
    if (!c) return HS_ERR_INVALID;
    int rc = drain_profile(c);
    if (rc != HS_OK) return rc;
    c->profile = on != 0;
    c->prof_launches = 0;
    c->prof_step_ms = 0.0;
    c->prof_reduce_ms = 0.0;
    return HS_OK;
}

int hs_profile_read(hs_ctx* c, uint64_t* launches, double* total_ms)
{
    if (!c || !launches || !total_ms) return HS_ERR_INVALID;
    int rc = drain_profile(c);
    if (rc != HS_OK) return rc;
    *launches = c->prof_launches;
    *total_ms = c->prof_step_ms + c->prof_reduce_ms;
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

}  // extern "C"
