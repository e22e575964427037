// hs_server.cpp - served simulation jobs over the batched step.
//
// The reference steps a single groom once per rendered frame from one host
// thread (src/main.cpp:257-285, src/Renderer.cpp:2309-2319); it has no job
// model.  This file adds one on top of the same step: jobs (a groom, a step
// count and a keyframed wind/sphere trajectory) are admitted while the HBM
// budget allows and stay resident in one device arena.  Each batched launch
// takes resident jobs in round-robin order up to the per-launch point limit
// and advances each of them by one step, so jobs that arrive while others
// run join a following launch (continuous batching).  Each job's result
// equals stepping it alone with hs_step: the batched kernels compute every
// groom exactly as the single-groom ones do, and a job's steps run in order
// on one stream.
//
// Threads: admission (FIFO; uploads on its own stream; each admitted job goes
// to the step loop holding the fewest points), the step loops (each owns a
// context and its stream; a job stays in its loop) and completion (reads
// finished jobs back on another stream while later launches run).  A loop
// builds its next launch on the host while its current one runs and keeps at
// most `depth` launches queued on its stream; the loops' launches run
// concurrently, so one launch's last round of workgroups and a loop's gaps
// between launches are filled by the others' work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hs_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// First-fit allocator over one device region; blocks of 256-byte multiples.
class Arena : public hsi::DevHeap {
public:
    hipError_t init(uint64_t bytes)
    {
        cap_ = hsi::round_up(std::max<uint64_t>(bytes, 256), 256);
        hipError_t e = hipMalloc((void**)&base_, cap_);
        if (e != hipSuccess) return e;
        free_[0] = cap_;
        return hipSuccess;
    }
    ~Arena() override
    {
        if (base_) (void)hipFree(base_);
    }
    hipError_t alloc(void** p, size_t bytes) override
    {
        std::lock_guard<std::mutex> lk(m_);
        const uint64_t need = hsi::round_up(std::max<size_t>(bytes, 1), 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const uint64_t off = it->first, size = it->second;
            free_.erase(it);
            if (size > need) free_[off + need] = size - need;
            used_[off] = need;
            *p = base_ + off;
            return hipSuccess;
        }
        return hipErrorOutOfMemory;
    }
    void release(void* p) override
    {
        std::lock_guard<std::mutex> lk(m_);
        const uint64_t off = (uint64_t)(static_cast<uint8_t*>(p) - base_);
        auto u = used_.find(off);
        if (u == used_.end()) return;
        uint64_t o = off, size = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(o);
        if (next != free_.end() && o + size == next->first) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == o) {
                o = prev->first;
                size += prev->second;
                free_.erase(prev);
            }
        }
        free_[o] = size;
    }

private:
    std::mutex m_;
    uint8_t* base_ = nullptr;
    uint64_t cap_ = 0;
    std::map<uint64_t, uint64_t> free_, used_;
};

enum class JobState { Queued, Admitted, Active, Reading, Done, Cancelled, Failed };

struct Job {
    uint64_t id = 0;
    hs_job spec{};
    uint64_t bytes = 0;   // groom block + read-back staging (held while resident)
    uint64_t upload = 0;  // upload staging, held only while the groom is built
    uint64_t points = 0;  // S*N
    Clock::time_point t_submit;
    uint32_t done_steps = 0;
    uint32_t groom_id = 0;
    struct Loop* loop = nullptr;  // the step loop (context) the job's groom lives in
    hsi::Groom g;
    void* staging = nullptr;
    JobState state = JobState::Queued;
};

void sphere_collider(const float c[3], hs_collider* out)
{
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < 4; ++k) {
        out->transform[k * 5] = 1.0f;
        out->inv[k * 5] = 1.0f;
        out->invTrans[k * 5] = 1.0f;
    }
    for (int k = 0; k < 3; ++k) {
        out->transform[12 + k] = c[k];
        out->inv[12 + k] = -c[k];
        out->invTrans[3 + 4 * k] = -c[k];  // transpose(inv): row 3 of the columns
    }
}

double weighted_percentile(std::vector<std::pair<double, uint64_t>> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    uint64_t total = 0;
    for (auto& e : v) total += e.second;
    const double target = q * (double)total;
    uint64_t acc = 0;
    for (auto& e : v) {
        acc += e.second;
        if ((double)acc >= target) return e.first;
    }
    return v.back().first;
}

// One batched launch on the stream: timing events around it, when the host
// enqueued it, what it stepped and the jobs whose last step it holds.
struct Flight {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    Clock::time_point enq;
    uint32_t jobs = 0;
    uint64_t points = 0;
    std::vector<Job*> last;
};

// One step loop: a context (its own stream, argument blob and pinned slots)
// and the thread that launches batched steps of the jobs assigned to it.
// Loops run their launches concurrently on their own streams.
constexpr int kMaxDepth = 2;  // launches queued per stream at once (the pinned argument slots)
struct Loop {
    hs_ctx* ctx = nullptr;
    std::vector<Job*> ready;  // uploaded, joins a following launch
    uint64_t points = 0;      // points of the jobs assigned to this loop
    Flight flight[kMaxDepth];
    std::thread t;
};

}  // namespace

struct hs_server {
    hs_server_config cfg{};
    int depth = 1;  // launches queued per loop
    std::vector<std::unique_ptr<Loop>> loops;
    Arena arena;
    hipStream_t up = nullptr, down = nullptr;

    std::mutex m;
    std::condition_variable cv_admit, cv_loop, cv_read, cv_done;
    std::deque<Job*> queue;     // submitted, not admitted
    std::deque<Job*> finished;  // last step done, to be read back
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    uint64_t next_id = 1;
    uint64_t resident = 0;       // bytes of admitted jobs (arena)
    uint64_t in_system = 0;      // admitted and not yet done
    Clock::time_point busy_since{};  // when in_system last became non-zero
    bool stop_admit = false, stop_loop = false, stop_read = false;
    std::thread t_admit, t_read;

    hs_server_stats st{};
    std::vector<std::pair<double, uint64_t>> step_ms;     // (launch device ms, jobs in it)
    std::vector<std::pair<double, uint64_t>> latency_ms;  // (enqueue -> completion ms, jobs in it)
    std::vector<std::pair<double, uint64_t>> wait_ms;
    int err_code = 0;
    std::string err_msg;

    void set_error(int code, const std::string& msg)
    {
        if (err_code == 0) {
            err_code = code;
            err_msg = msg;
        }
    }
    bool fits(const Job* j) const { return resident + j->bytes + j->upload <= cfg.hbm_budget; }
    // in_system changes (with `m` held); busy_s counts the time it is non-zero
    void enter()
    {
        if (in_system++ == 0) busy_since = Clock::now();
    }
    void leave()
    {
        if (--in_system == 0) st.busy_s += seconds_since(busy_since);
    }
    void fail_job(Job* j);
    void admit_loop();
    void step_loop(Loop& L);
    void read_loop();
};

// A resident job whose device state is unknown after a failed launch: its
// memory goes back to the arena (the caller has drained the stream) and
// waiters see the error.  Called with `m` held.
void hs_server::fail_job(Job* j)
{
    hsi::Groom g;
    if (j->loop && hsi::ctx_take(j->loop->ctx, j->groom_id, &g)) hsi::groom_release(g);
    hsi::groom_release(j->g);  // a job whose last step was enqueued holds its groom itself
    if (j->loop) j->loop->points -= j->points;
    if (j->staging) arena.release(j->staging);
    j->staging = nullptr;
    resident -= j->bytes;
    leave();
    j->state = JobState::Failed;
}

void hs_server::admit_loop()
{
    (void)hipSetDevice(cfg.device);
    std::unique_lock<std::mutex> lk(m);
    while (true) {
        cv_admit.wait(lk, [&] { return stop_admit || (!queue.empty() && fits(queue.front())); });
        if (stop_admit) break;
        Job* j = queue.front();
        queue.pop_front();
        resident += j->bytes;
        enter();
        st.admitted += 1;
        st.peak_resident_bytes = std::max(st.peak_resident_bytes, resident);
        j->state = JobState::Admitted;
        lk.unlock();

        std::string why;
        int rc = hsi::groom_build(&j->spec.groom, j->spec.strands_aos, up, &arena, &j->g, &why);
        hipError_t e = hipSuccess;
        if (rc == HS_OK && j->spec.result_aos) {
            e = arena.alloc(&j->staging, (size_t)j->points * 48);
            if (e != hipSuccess) hsi::groom_release(j->g);
        }
        lk.lock();
        if (rc == HS_ERR_OUT_OF_MEMORY || e == hipErrorOutOfMemory) {
            // the arena is fragmented: wait for a release unless nothing else holds memory
            resident -= j->bytes;
            leave();
            st.admitted -= 1;
            if (resident == 0) {
                j->state = JobState::Failed;
                set_error(HS_ERR_OUT_OF_MEMORY, "hs_server: job does not fit the arena");
                cv_done.notify_all();
            } else {
                j->state = JobState::Queued;
                queue.push_front(j);
                cv_admit.wait(lk);  // until a completion or a new job changes the picture
            }
            continue;
        }
        if (rc != HS_OK || e != hipSuccess) {
            resident -= j->bytes;
            leave();
            j->state = JobState::Failed;
            set_error(rc != HS_OK ? rc : HS_ERR_HIP, "hs_server admission: " + why);
            cv_done.notify_all();
            continue;
        }
        // the loop with the fewest points takes the job
        Loop* best = loops[0].get();
        for (auto& L : loops)
            if (L->points < best->points) best = L.get();
        j->loop = best;
        best->points += j->points;
        best->ready.push_back(j);
        cv_loop.notify_all();
    }
}

void hs_server::step_loop(Loop& L)
{
    (void)hipSetDevice(cfg.device);
    std::deque<Job*> active;  // resident jobs in round-robin order
    std::vector<Job*> batch, skipped;
    std::vector<hs_batch_item> items;
    std::vector<hs_collider> cols;
    std::vector<hs_step_params> params;
    hs_ctx* ctx = L.ctx;
    Flight* flight = L.flight;
    int head = 0, inflight = 0;

    // After a failed enqueue or launch: drain the stream, then fail every
    // resident job (those stepped in the failed launch included).
    auto fail_all = [&](int code, const std::string& msg) {
        (void)hipStreamSynchronize(ctx->stream);
        std::lock_guard<std::mutex> lk(m);
        set_error(code, msg);
        for (int q = 0; q < inflight; ++q) {
            Flight& f = flight[(head + q) % depth];
            for (Job* j : f.last) fail_job(j);
            f.last.clear();
        }
        inflight = 0;
        for (Job* j : batch) fail_job(j);
        for (Job* j : active) fail_job(j);
        batch.clear();
        active.clear();
        cv_done.notify_all();
        cv_admit.notify_all();
    };
    // Wait for the oldest launch on the stream, record its times and hand
    // the jobs it finished to the read-back thread.
    auto retire = [&]() -> bool {
        Flight& f = flight[head];
        hipError_t e = hipEventSynchronize(f.t1);
        const Clock::time_point seen = Clock::now();
        float dev_ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&dev_ms, f.t0, f.t1);
        if (e != hipSuccess) {
            fail_all(HS_ERR_HIP, std::string("hs_server step: ") + hipGetErrorString(e));
            return false;
        }
        head = (head + 1) % depth;
        --inflight;
        std::lock_guard<std::mutex> lk(m);
        st.launches += 1;
        st.job_steps += f.jobs;
        st.point_updates += f.points;
        st.gpu_busy_s += dev_ms * 1e-3;
        step_ms.emplace_back(dev_ms, f.jobs);
        latency_ms.emplace_back(std::chrono::duration<double, std::milli>(seen - f.enq).count(), f.jobs);
        for (Job* j : f.last) {
            j->state = JobState::Reading;
            finished.push_back(j);
        }
        if (!f.last.empty()) cv_read.notify_one();
        f.last.clear();
        return true;
    };

    while (true) {
        {
            std::unique_lock<std::mutex> lk(m);
            if (active.empty() && inflight == 0) {
                cv_loop.wait(lk, [&] { return stop_loop || !L.ready.empty(); });
                if (L.ready.empty()) break;  // stop_loop with nothing left
            }
            for (Job* j : L.ready) {
                j->groom_id = hsi::ctx_adopt(ctx, std::move(j->g));
                j->state = JobState::Active;
                active.push_back(j);
            }
            L.ready.clear();
        }
        if (active.empty()) {
            retire();  // only launches on the stream are left
            continue;
        }
        // This launch: jobs from the front of the round-robin order while
        // they fit the point limit.  A job that does not fit keeps its place
        // at the front, so it leads a following launch (every job fits one
        // launch alone: submit checks that).
        batch.clear();
        skipped.clear();
        uint64_t pts = 0;
        for (size_t scanned = 0; !active.empty() && scanned < 4096 && pts < cfg.max_points_per_step; ++scanned) {
            Job* j = active.front();
            active.pop_front();
            if (pts + j->points <= cfg.max_points_per_step) {
                batch.push_back(j);
                pts += j->points;
            } else {
                skipped.push_back(j);
            }
        }
        for (auto it = skipped.rbegin(); it != skipped.rend(); ++it) active.push_front(*it);
        const size_t n = batch.size();
        items.resize(n);
        cols.resize(n);
        params.resize(n);
        for (size_t k = 0; k < n; ++k) {
            Job* j = batch[k];
            const float t = (float)(j->done_steps + 1) * j->spec.dt;
            float wind[3], sph[3];
            hs_keyframes_eval(j->spec.keys, j->spec.num_keys, t, wind, sph);
            params[k] = j->spec.params;
            for (int c = 0; c < 3; ++c) params[k].wind[c] = wind[c];
            sphere_collider(sph, &cols[k]);
            items[k].groom_id = j->groom_id;
            items[k].time.deltaTime = j->spec.dt;
            items[k].time.totalTime = t;
            items[k].colliders = &cols[k];
            items[k].num_colliders = params[k].num_colliders > 0 ? 1u : 0u;
            items[k].params = &params[k];
        }
        // the host work above overlapped the launches on the stream; now
        // make room for this one
        if (inflight == depth && !retire()) continue;
        Flight& f = flight[(head + inflight) % depth];
        f.enq = Clock::now();
        int rc = HS_OK;
        hipError_t e = hipEventRecord(f.t0, ctx->stream);
        if (e == hipSuccess) rc = hs_step_batch(ctx, items.data(), (uint32_t)n);
        if (e == hipSuccess && rc == HS_OK) e = hipEventRecord(f.t1, ctx->stream);
        if (rc != HS_OK || e != hipSuccess) {
            // the device state of the resident jobs is unknown: fail them
            fail_all(rc != HS_OK ? rc : HS_ERR_HIP, rc != HS_OK ? std::string("hs_server step: ") + ctx->err
                                                                  : std::string("hs_server step: ") +
                                                                        hipGetErrorString(e));
            continue;
        }
        ++inflight;
        f.jobs = (uint32_t)n;
        f.points = pts;
        std::lock_guard<std::mutex> lk(m);
        st.peak_jobs_per_launch = std::max<uint32_t>(st.peak_jobs_per_launch, (uint32_t)n);
        for (Job* j : batch) {
            if (j->done_steps == 0)
                wait_ms.emplace_back(std::chrono::duration<double, std::milli>(f.enq - j->t_submit).count(), 1);
            j->done_steps += 1;
            if (j->done_steps >= j->spec.steps) {
                // out of the context now; its memory stays until read back
                hsi::ctx_take(ctx, j->groom_id, &j->g);
                L.points -= j->points;
                j->loop = nullptr;
                f.last.push_back(j);
            } else {
                active.push_back(j);
            }
        }
        batch.clear();
    }
}

void hs_server::read_loop()
{
    (void)hipSetDevice(cfg.device);
    std::unique_lock<std::mutex> lk(m);
    while (true) {
        cv_read.wait(lk, [&] { return stop_read || !finished.empty(); });
        if (finished.empty()) break;
        Job* j = finished.front();
        finished.pop_front();
        lk.unlock();
        // the step loop saw the job's last launch complete: the state is final
        hipError_t e = hipSuccess;
        if (j->spec.result_aos) {
            e = hsi::groom_unpack(j->g, static_cast<float*>(j->staging), down);
            if (e == hipSuccess)
                e = hipMemcpyAsync(j->spec.result_aos, j->staging, (size_t)j->points * 48, hipMemcpyDeviceToHost,
                                   down);
            if (e == hipSuccess) e = hipStreamSynchronize(down);
        }
        hsi::groom_release(j->g);
        if (j->staging) arena.release(j->staging);
        j->staging = nullptr;
        lk.lock();
        resident -= j->bytes;
        leave();
        if (e != hipSuccess) {
            set_error(HS_ERR_HIP, std::string("hs_server read-back: ") + hipGetErrorString(e));
            j->state = JobState::Failed;
        } else {
            j->state = JobState::Done;
            st.completed += 1;
        }
        cv_done.notify_all();
        cv_admit.notify_all();
    }
}

extern "C" {

void hs_keyframes_eval(const hs_keyframe* keys, uint32_t n, float t, float wind[3], float sphere[3])
{
    if (!keys || n == 0) {
        for (int c = 0; c < 3; ++c) wind[c] = sphere[c] = 0.0f;
        return;
    }
    const hs_keyframe* a = &keys[0];
    const hs_keyframe* b = &keys[0];
    if (t >= keys[n - 1].t) {
        a = b = &keys[n - 1];
    } else if (t > keys[0].t) {
        uint32_t i = 0;
        while (i + 1 < n && keys[i + 1].t <= t) ++i;
        a = &keys[i];
        b = &keys[i + 1];
    }
    if (a == b || !(b->t > a->t)) {
        for (int c = 0; c < 3; ++c) {
            wind[c] = a->wind[c];
            sphere[c] = a->sphere[c];
        }
        return;
    }
    const float s = (t - a->t) / (b->t - a->t);
    for (int c = 0; c < 3; ++c) {
        wind[c] = a->wind[c] + s * (b->wind[c] - a->wind[c]);
        sphere[c] = a->sphere[c] + s * (b->sphere[c] - a->sphere[c]);
    }
}

// Releases what hs_server_create made (the threads are not running).
static void server_free(hs_server* s)
{
    (void)hipSetDevice(s->cfg.device);
    if (s->up) (void)hipStreamSynchronize(s->up);
    if (s->down) (void)hipStreamSynchronize(s->down);
    for (auto& L : s->loops) {
        if (L->ctx) hs_destroy(L->ctx);
        for (Flight& f : L->flight) {
            if (f.t0) (void)hipEventDestroy(f.t0);
            if (f.t1) (void)hipEventDestroy(f.t1);
        }
    }
    if (s->up) (void)hipStreamDestroy(s->up);
    if (s->down) (void)hipStreamDestroy(s->down);
    delete s;
}

int hs_server_create(const hs_server_config* cfg, hs_server** out)
{
    if (!cfg || !out) return HS_ERR_INVALID;
    *out = nullptr;
    if (cfg->hbm_budget == 0 || cfg->max_points_per_step == 0) return HS_ERR_INVALID;
    if (cfg->step_loops > HS_SERVER_MAX_STEP_LOOPS || cfg->launch_depth > 2) return HS_ERR_INVALID;
    hs_server* s = new (std::nothrow) hs_server();
    if (!s) return HS_ERR_OUT_OF_MEMORY;
    s->cfg = *cfg;
    s->depth = cfg->launch_depth ? (int)cfg->launch_depth : 1;
    const hs_tile zero{};
    const bool given = std::memcmp(&cfg->tile, &zero, sizeof(zero)) != 0;
    const hs_tile tile = given ? cfg->tile : hs_tile{2048, 2048, 0, 256, 8};
    const uint32_t nloops = cfg->step_loops ? cfg->step_loops : 4;
    int rc = HS_OK;
    hipError_t e = hipSuccess;
    for (uint32_t k = 0; k < nloops && rc == HS_OK && e == hipSuccess; ++k) {
        s->loops.push_back(std::make_unique<Loop>());
        Loop& L = *s->loops.back();
        rc = hs_create(&L.ctx, cfg->device);
        if (rc == HS_OK && cfg->use_graphs) rc = hs_set_batch_graphs(L.ctx, 1);
        if (rc == HS_OK) rc = hs_set_tile(L.ctx, &tile);
        for (Flight& f : L.flight) {
            if (rc == HS_OK && e == hipSuccess) e = hipEventCreate(&f.t0);
            if (rc == HS_OK && e == hipSuccess) e = hipEventCreate(&f.t1);
        }
    }
    if (rc == HS_OK && e == hipSuccess) e = s->arena.init(cfg->hbm_budget);
    if (rc == HS_OK && e == hipSuccess) e = hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking);
    if (rc == HS_OK && e == hipSuccess) e = hipStreamCreateWithFlags(&s->down, hipStreamNonBlocking);
    if (rc != HS_OK || e != hipSuccess) {
        server_free(s);
        if (rc != HS_OK) return rc;
        return e == hipErrorOutOfMemory ? HS_ERR_OUT_OF_MEMORY : HS_ERR_HIP;
    }
    s->t_admit = std::thread([s] { s->admit_loop(); });
    for (auto& L : s->loops) {
        Loop* l = L.get();
        l->t = std::thread([s, l] { s->step_loop(*l); });
    }
    s->t_read = std::thread([s] { s->read_loop(); });
    *out = s;
    return HS_OK;
}

int hs_server_submit(hs_server* s, const hs_job* job, uint64_t* job_id)
{
    if (!s || !job) return HS_ERR_INVALID;
    std::string why;
    int rc = hsi::check_groom_desc(&job->groom, job->strands_aos, &why);
    if (rc != HS_OK) return rc;
    // everything hs_step_batch would reject is rejected here: a job that
    // failed inside a batched launch would take the launch's other jobs with it
    if (job->steps == 0 || !(job->dt > 0.0f) || !std::isfinite(job->dt) || !job->keys || job->num_keys == 0 ||
        job->params.num_colliders > 1 || job->params.mesh_id != 0 || job->params.forcing > 2)
        return HS_ERR_INVALID;
    auto j = std::make_unique<Job>();
    j->spec = *job;
    j->points = (uint64_t)job->groom.num_strands * job->groom.num_points;
    j->bytes = hs_groom_bytes(job->groom.num_strands, job->groom.num_points, job->groom.grid_dim,
                              job->groom.flags) +
               (job->result_aos ? hsi::round_up(j->points * 48, 256) : 0);
    j->upload = hsi::round_up(j->points * 48, 256);
    if (j->bytes + j->upload > s->cfg.hbm_budget || j->points > s->cfg.max_points_per_step) return HS_ERR_BUDGET;
    j->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->m);
    j->id = s->next_id++;
    if (job_id) *job_id = j->id;
    s->queue.push_back(j.get());
    s->jobs[j->id] = std::move(j);
    s->st.submitted += 1;
    s->cv_admit.notify_one();
    return HS_OK;
}

int hs_server_wait(hs_server* s, uint64_t job_id, double timeout_s)
{
    if (!s) return HS_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return HS_ERR_NOT_FOUND;
    Job* j = it->second.get();
    auto finished = [&] {
        return j->state == JobState::Done || j->state == JobState::Cancelled || j->state == JobState::Failed;
    };
    if (timeout_s < 0.0)
        s->cv_done.wait(lk, finished);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished))
        return HS_ERR_STATE;
    if (j->state == JobState::Done) return HS_OK;
    return j->state == JobState::Cancelled ? HS_ERR_NOT_FOUND : (s->err_code ? s->err_code : HS_ERR_HIP);
}

int hs_server_cancel_pending(hs_server* s, uint64_t* cancelled)
{
    if (!s) return HS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    uint64_t n = 0;
    for (Job* j : s->queue) {
        j->state = JobState::Cancelled;
        ++n;
    }
    s->queue.clear();
    s->st.cancelled += n;
    if (cancelled) *cancelled = n;
    s->cv_done.notify_all();
    return HS_OK;
}

int hs_server_drain(hs_server* s)
{
    if (!s) return HS_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->in_system == 0; });
    return s->err_code;
}

int hs_server_get_stats(hs_server* s, hs_server_stats* out)
{
    if (!s || !out) return HS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->st;
    out->p50_step_ms = weighted_percentile(s->step_ms, 0.50);
    out->p99_step_ms = weighted_percentile(s->step_ms, 0.99);
    double mx = 0.0;
    for (auto& e : s->step_ms) mx = std::max(mx, e.first);
    out->max_step_ms = mx;
    out->p50_latency_ms = weighted_percentile(s->latency_ms, 0.50);
    out->p99_latency_ms = weighted_percentile(s->latency_ms, 0.99);
    out->p50_wait_ms = weighted_percentile(s->wait_ms, 0.50);
    out->p99_wait_ms = weighted_percentile(s->wait_ms, 0.99);
    return HS_OK;
}

int hs_server_error(hs_server* s, const char** message)
{
    if (!s) return HS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    if (message) *message = s->err_msg.c_str();
    return s->err_code;
}

void hs_server_destroy(hs_server* s)
{
    if (!s) return;
    hs_server_cancel_pending(s, nullptr);
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop_admit = true;
    }
    s->cv_admit.notify_all();
    s->t_admit.join();
    // admission has stopped: whatever it admitted is in a loop's `ready` or beyond
    {
        std::unique_lock<std::mutex> lk(s->m);
        s->cv_done.wait(lk, [&] { return s->in_system == 0; });
        s->stop_loop = true;
        s->stop_read = true;
    }
    s->cv_loop.notify_all();
    s->cv_read.notify_all();
    for (auto& L : s->loops) L->t.join();
    s->t_read.join();
    server_free(s);
}

}  // extern "C"
