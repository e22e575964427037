// hs_server.cpp - served simulation jobs over the batched step.
//
// The reference steps a single groom once per rendered frame from one host
// thread (src/main.cpp:257-285, src/Renderer.cpp:2309-2319); it has no job
// model.  This file adds one on top of the same step: jobs (a groom, a step
// count and a keyframed wind/sphere trajectory) are admitted while an HBM
// budget and a per-launch point limit allow, stay resident in one device
// arena, and every admitted job advances by one step in each batched launch,
// so jobs that arrive while others run join the next launch (continuous
// batching).  Each job's result equals stepping it alone with hs_step: the
// batched kernels compute every groom exactly as the single-groom ones do.
//
// Threads: admission (FIFO; uploads on its own stream), the step loop (owns
// the context and its stream) and completion (reads finished jobs back on a
// third stream while the next launch runs).
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

}  // namespace

struct hs_server {
    hs_server_config cfg{};
    hs_ctx* ctx = nullptr;
    Arena arena;
    hipStream_t up = nullptr, down = nullptr;
    hipEvent_t step_done = nullptr;

    std::mutex m;
    std::condition_variable cv_admit, cv_loop, cv_read, cv_done;
    std::deque<Job*> queue;     // submitted, not admitted
    std::vector<Job*> ready;    // uploaded, joins the next launch
    std::deque<Job*> finished;  // last step done, to be read back
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    uint64_t next_id = 1;
    uint64_t resident = 0;       // bytes of admitted jobs (arena)
    uint64_t admitted_points = 0;
    uint64_t in_system = 0;      // admitted and not yet done
    bool stop_admit = false, stop_loop = false, stop_read = false;
    std::thread t_admit, t_loop, t_read;

    hs_server_stats st{};
    std::vector<std::pair<double, uint64_t>> step_ms;  // (launch ms, jobs in it)
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
    bool fits(const Job* j) const
    {
        return resident + j->bytes + j->upload <= cfg.hbm_budget &&
               admitted_points + j->points <= cfg.max_points_per_step;
    }
    void admit_loop();
    void step_loop();
    void read_loop();
};

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
        admitted_points += j->points;
        in_system += 1;
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
            admitted_points -= j->points;
            in_system -= 1;
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
            admitted_points -= j->points;
            in_system -= 1;
            j->state = JobState::Failed;
            set_error(rc != HS_OK ? rc : HS_ERR_HIP, "hs_server admission: " + why);
            cv_done.notify_all();
            continue;
        }
        ready.push_back(j);
        cv_loop.notify_one();
    }
}

void hs_server::step_loop()
{
    (void)hipSetDevice(cfg.device);
    std::vector<Job*> active;
    std::vector<hs_batch_item> items;
    std::vector<hs_collider> cols;
    std::vector<hs_step_params> params;
    Clock::time_point busy_since{};
    bool busy = false;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(m);
            if (active.empty()) {
                if (busy) {
                    st.busy_s += seconds_since(busy_since);
                    busy = false;
                }
                cv_loop.wait(lk, [&] { return stop_loop || !ready.empty(); });
                if (ready.empty()) break;  // stop_loop with nothing left
            }
            for (Job* j : ready) {
                j->groom_id = hsi::ctx_adopt(ctx, std::move(j->g));
                j->state = JobState::Active;
                active.push_back(j);
            }
            ready.clear();
        }
        if (!busy) {
            busy = true;
            busy_since = Clock::now();
        }
        const size_t n = active.size();
        items.resize(n);
        cols.resize(n);
        params.resize(n);
        uint64_t pts = 0;
        for (size_t k = 0; k < n; ++k) {
            Job* j = active[k];
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
            pts += j->points;
        }
        const Clock::time_point t0 = Clock::now();
        int rc = hs_step_batch(ctx, items.data(), (uint32_t)n);
        hipError_t e = hipSuccess;
        if (rc == HS_OK) {
            e = hipEventRecord(step_done, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(step_done);
        }
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        std::lock_guard<std::mutex> lk(m);
        if (rc != HS_OK || e != hipSuccess) {
            set_error(rc != HS_OK ? rc : HS_ERR_HIP,
                      rc != HS_OK ? std::string("hs_server step: ") + ctx->err
                                  : std::string("hs_server step: ") + hipGetErrorString(e));
            // the device state of these jobs is unknown: fail them
            for (Job* j : active) {
                hsi::Groom g;
                if (hsi::ctx_take(ctx, j->groom_id, &g)) hsi::groom_release(g);
                if (j->staging) arena.release(j->staging);
                resident -= j->bytes;
                admitted_points -= j->points;
                in_system -= 1;
                j->state = JobState::Failed;
            }
            active.clear();
            cv_done.notify_all();
            cv_admit.notify_all();
            continue;
        }
        st.launches += 1;
        st.job_steps += n;
        st.point_updates += pts;
        st.peak_jobs_per_launch = std::max<uint32_t>(st.peak_jobs_per_launch, (uint32_t)n);
        step_ms.emplace_back(ms, n);
        size_t keep = 0;
        bool freed = false;
        for (size_t k = 0; k < n; ++k) {
            Job* j = active[k];
            if (j->done_steps == 0)
                wait_ms.emplace_back(std::chrono::duration<double, std::milli>(t0 - j->t_submit).count(), 1);
            j->done_steps += 1;
            if (j->done_steps >= j->spec.steps) {
                hsi::ctx_take(ctx, j->groom_id, &j->g);
                j->state = JobState::Reading;
                admitted_points -= j->points;
                finished.push_back(j);
                freed = true;
            } else {
                active[keep++] = j;
            }
        }
        active.resize(keep);
        if (freed) {
            cv_read.notify_one();
            cv_admit.notify_all();
        }
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
        // the step loop waited for the last launch: the state is final
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
        in_system -= 1;
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

int hs_server_create(const hs_server_config* cfg, hs_server** out)
{
    if (!cfg || !out) return HS_ERR_INVALID;
    *out = nullptr;
    if (cfg->hbm_budget == 0 || cfg->max_points_per_step == 0) return HS_ERR_INVALID;
    std::unique_ptr<hs_server> s(new (std::nothrow) hs_server());
    if (!s) return HS_ERR_OUT_OF_MEMORY;
    s->cfg = *cfg;
    int rc = hs_create(&s->ctx, cfg->device);
    if (rc != HS_OK) return rc;
    if (cfg->use_graphs) hs_set_batch_graphs(s->ctx, 1);
    hipError_t e = s->arena.init(cfg->hbm_budget);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->down, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->step_done, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (s->up) (void)hipStreamDestroy(s->up);
        if (s->down) (void)hipStreamDestroy(s->down);
        if (s->step_done) (void)hipEventDestroy(s->step_done);
        hs_destroy(s->ctx);
        return e == hipErrorOutOfMemory ? HS_ERR_OUT_OF_MEMORY : HS_ERR_HIP;
    }
    hs_server* p = s.get();
    p->t_admit = std::thread([p] { p->admit_loop(); });
    p->t_loop = std::thread([p] { p->step_loop(); });
    p->t_read = std::thread([p] { p->read_loop(); });
    *out = s.release();
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
    // admission has stopped: whatever it admitted is in `ready` or beyond
    {
        std::unique_lock<std::mutex> lk(s->m);
        s->cv_done.wait(lk, [&] { return s->in_system == 0; });
        s->stop_loop = true;
        s->stop_read = true;
    }
    s->cv_loop.notify_all();
    s->cv_read.notify_all();
    s->t_loop.join();
    s->t_read.join();
    (void)hipSetDevice(s->cfg.device);
    (void)hipStreamSynchronize(s->up);
    (void)hipStreamSynchronize(s->down);
    hs_destroy(s->ctx);
    (void)hipStreamDestroy(s->up);
    (void)hipStreamDestroy(s->down);
    (void)hipEventDestroy(s->step_done);
    delete s;
}

}  // extern "C"
