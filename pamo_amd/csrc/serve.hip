// Batched and served forms of the PaMO pipeline on one GPU.
//
// A pool of slots, each owning a stage-2 context and (when stage 3 is on) a
// stage-3 context -- each with its own HIP stream -- and one worker thread.
// Jobs (one mesh each: centring, the stage-2 round loop of PaMO.run, and
// stage 3's five steps against the mesh itself, simp_cuda/pamo/__init__.py:
// 122-205) are admitted in arrival order when a slot is free and the
// estimated HBM footprint of the running jobs plus the new one fits the
// budget; admitted jobs run concurrently, so the short, latency-bound
// kernels of the late rounds of one mesh overlap with the work of others.
// A job's result does not depend on what else runs: every kernel of a job
// runs on its slot's stream and contexts hold no shared state.
//
// pamo_batch_run packs a batch in a CSR mesh-of-meshes layout (concatenated
// vertices and faces with per-mesh offsets; face indices local to their
// mesh) and submits the meshes longest first.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hip_common.h"
#include "pamo_hip.h"

namespace pamo {
namespace serve {

constexpr int kB = 256;

__global__ void shift_k(const float* __restrict__ v, long long n3, float mx, float my, float mz, float sign,
                        float* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const float m = (i % 3 == 0) ? mx : ((i % 3 == 1) ? my : mz);
    // sign -1: v - m (centring, pamo/__init__.py:144); +1: v + m (:187)
    out[i] = sign < 0.0f ? v[i] - m : v[i] + m;
}

struct Job {
    long long id = 0;
    int index = -1;  // position in a batch (-1: served job)
    const float* dv = nullptr;
    const int* df = nullptr;
    int nv = 0, nf = 0;
    double ratio = 0.01;
    int stage3 = 0, tolerance = 4;
    float threshold = 1e-3f;
    long long iters = 100000, min_verts = 0;
    long long est = 0;
    // results
    float* out_v = nullptr;  // device, owned by the caller (batch) or by the job (serve)
    int* out_f = nullptr;
    bool owns_out = false;
    int out_nv = 0, out_nf = 0, rounds = 0;
    int status = 0;
    std::string err;
    bool done = false;
    double t_submit = 0, t_start = 0, t_end = 0;
};

struct Slot {
    pamo_ctx* c2 = nullptr;
    pamo_s3* c3 = nullptr;
    hipStream_t st = nullptr;
    float *vbuf = nullptr, *obuf = nullptr;
    int* fbuf = nullptr;
    size_t cap_v = 0, cap_f = 0;
    std::thread th;
};

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// HBM footprint of one job: the stage-2 context's arrays (about 231 B per
// face and 128 B per vertex over the round, with the LBVH) plus, with stage
// 3, the input-surface tree and the CG state of the output mesh.
static long long estimate_bytes(int nv, int nf, int stage3) {
    long long b = 260LL * nf + 160LL * nv + (12LL * nv + 12LL * nf);
    if (stage3) b += 160LL * nf + 64LL * nv + (long long)(0.01 * nf) * 4096;
    return b;
}

struct Engine {
    int device = 0;
    bool has_s3 = false;
    pamo_s3_params s3p{};
    std::vector<Slot> slots;
    hipStream_t io = nullptr;  // result copies of pamo_serve_fetch
    pamo_batch_done_fn on_done = nullptr;
    void* on_done_user = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::unordered_map<long long, Job*> jobs;
    long long next_id = 1, budget = 0, in_use = 0;
    int running = 0;
    bool stop = false;
};

// Grows a slot's staging buffers.  The old buffers are released first (the
// slot's work on them has finished); a failed allocation leaves the slot
// with no buffer and zero capacity, so the next job allocates again instead
// of writing through a stale pointer.
static int ensure_slot_buffers(Slot& s, int nv, int nf) {
    if ((size_t)nv > s.cap_v) {
        if (s.vbuf) (void)hipFree(s.vbuf);
        if (s.obuf) (void)hipFree(s.obuf);
        s.vbuf = s.obuf = nullptr;
        s.cap_v = 0;
        const size_t cap = (size_t)nv + nv / 4 + 64;
        float *vb = nullptr, *ob = nullptr;
        hipError_t e = hipMalloc((void**)&vb, 12 * cap);
        if (e == hipSuccess) e = hipMalloc((void**)&ob, 12 * cap);
        if (e != hipSuccess) {
            if (vb) (void)hipFree(vb);
            PAMO_HIP(e);
        }
        s.vbuf = vb;
        s.obuf = ob;
        s.cap_v = cap;
    }
    if ((size_t)nf > s.cap_f) {
        if (s.fbuf) (void)hipFree(s.fbuf);
        s.fbuf = nullptr;
        s.cap_f = 0;
        const size_t cap = (size_t)nf + nf / 4 + 64;
        int* fb = nullptr;
        PAMO_HIP(hipMalloc((void**)&fb, 12 * cap));
        s.fbuf = fb;
        s.cap_f = cap;
    }
    return 0;
}

#define SV_TRY(call)                            \
    do {                                        \
        int rc_ = (call);                       \
        if (rc_ != 0) {                         \
            j.err = ::pamo::get_error();        \
            return rc_;                         \
        }                                       \
    } while (0)

// One mesh through centring, stage 2 and (optionally) stage 3, on one slot.
static int run_job(Engine& e, Slot& s, Job& j) {
    SV_TRY(ensure_slot_buffers(s, j.nv, j.nf));
    // each job behaves as a fresh PaMO object, whatever ran on this slot before
    SV_TRY(pamo_ctx_reset(s.c2));
    float mean[3];
    SV_TRY(pamo_centroid(s.c2, j.dv, j.nv, j.df, j.nf, mean));
    const long long n3 = 3LL * j.nv;
    hipLaunchKernelGGL(shift_k, dim3(grid_for(n3, kB)), dim3(kB), 0, s.st, j.dv, n3, mean[0], mean[1], mean[2], -1.0f,
                       s.vbuf);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s.st) != hipSuccess) {
        j.err = "centring failed";
        return 1;
    }
    int onv = 0, onf = 0, rounds = 0;
    SV_TRY(pamo_simplify(s.c2, s.vbuf, j.nv, j.df, j.nf, j.ratio, j.tolerance, j.threshold, j.iters, j.min_verts,
                         s.obuf, s.fbuf, &onv, &onf, nullptr, 0, &rounds));
    // + centring offset (float32, as the reference adds tris_mean)
    const long long o3 = 3LL * onv;
    hipLaunchKernelGGL(shift_k, dim3(grid_for(std::max(o3, 1LL), kB)), dim3(kB), 0, s.st, s.obuf, o3, mean[0],
                       mean[1], mean[2], 1.0f, j.stage3 ? s.vbuf : j.out_v);
    if (hipGetLastError() != hipSuccess) {
        j.err = "offset failed";
        return 1;
    }
    if (j.stage3 && onf == 0) {
        if (hipMemcpyAsync(j.out_v, s.vbuf, 12 * (size_t)onv, hipMemcpyDeviceToDevice, s.st) != hipSuccess) {
            j.err = "output copy failed";
            return 1;
        }
    } else if (j.stage3) {
        if (hipStreamSynchronize(s.st) != hipSuccess) {
            j.err = "stream sync failed";
            return 1;
        }
        double scale = 1.0, t[3];
        SV_TRY(pamo_s3_normalization(s.c3, j.dv, j.nv, &scale, t));
        SV_TRY(pamo_s3_set_target(s.c3, j.dv, j.nv, j.df, j.nf, scale, t));
        // trimesh.Trimesh(process=True)'s vertex merge of the stage-2 result
        SV_TRY(pamo_s3_merge_vertices(s.c3, s.vbuf, onv, s.fbuf, onf, s.obuf, &onv));
        SV_TRY(pamo_s3_set_mesh_device(s.c3, s.obuf, onv, s.fbuf, onf, scale, t));
        SV_TRY(pamo_s3_step(s.c3, 5));
        SV_TRY(pamo_s3_get_vertices(s.c3, j.out_v, scale, t));
    }
    if (hipMemcpyAsync(j.out_f, s.fbuf, 12 * (size_t)onf, hipMemcpyDeviceToDevice, s.st) != hipSuccess ||
        hipStreamSynchronize(s.st) != hipSuccess) {
        j.err = "output copy failed";
        return 1;
    }
    j.out_nv = onv;
    j.out_nf = onf;
    j.rounds = rounds;
    (void)e;
    return 0;
}

static void worker(Engine* e, int si) {
    Slot& s = e->slots[si];
    (void)hipSetDevice(e->device);
    while (true) {
        Job* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(e->mu);
            e->cv_work.wait(lk, [&] {
                if (e->stop) return true;
                if (e->queue.empty()) return false;
                // arrival order; a job larger than the whole budget runs alone
                Job* h = e->queue.front();
                return e->in_use + h->est <= e->budget || e->running == 0;
            });
            if (e->stop && e->queue.empty()) return;
            if (e->stop) return;
            j = e->queue.front();
            e->queue.pop_front();
            e->in_use += j->est;
            ++e->running;
        }
        j->t_start = now_s();
        pamo::clear_error();
        j->status = run_job(*e, s, *j);
        j->t_end = now_s();
        {
            std::lock_guard<std::mutex> lk(e->mu);
            e->in_use -= j->est;
            --e->running;
            j->done = true;
        }
        if (e->on_done && j->index >= 0) e->on_done(j->index, j->status, e->on_done_user);
        e->cv_done.notify_all();
        e->cv_work.notify_all();
    }
}

static void destroy(Engine* e) {
    {
        std::lock_guard<std::mutex> lk(e->mu);
        e->stop = true;
    }
    e->cv_work.notify_all();
    for (Slot& s : e->slots)
        if (s.th.joinable()) s.th.join();
    (void)hipSetDevice(e->device);
    if (e->io) (void)hipStreamDestroy(e->io);
    for (Slot& s : e->slots) {
        if (s.c2) pamo_ctx_destroy(s.c2);
        if (s.c3) pamo_s3_destroy(s.c3);
        if (s.st) (void)hipStreamDestroy(s.st);
        if (s.vbuf) (void)hipFree(s.vbuf);
        if (s.obuf) (void)hipFree(s.obuf);
        if (s.fbuf) (void)hipFree(s.fbuf);
    }
    for (auto& kv : e->jobs) {
        Job* j = kv.second;
        if (j->owns_out) {
            if (j->out_v) (void)hipFree(j->out_v);
            if (j->out_f) (void)hipFree(j->out_f);
        }
        delete j;
    }
    delete e;
}

static int create(Engine** out, int device, int n_slots, long long budget, const pamo_s3_params* s3,
                  int exact_guard) {
    PAMO_CHECK(out && n_slots > 0 && n_slots <= 64, "pamo serve/batch: 1..64 slots");
    PAMO_CHECK(exact_guard >= 0 && exact_guard <= 2, "pamo serve/batch: exact_guard must be 0, 1 or 2");
    Engine* e = new Engine();
    e->device = device;
    e->budget = budget > 0 ? budget : (long long)200 << 30;
    e->has_s3 = s3 != nullptr;
    if (s3) e->s3p = *s3;
    e->slots.resize(n_slots);
    PAMO_HIP(hipSetDevice(device));
    if (hipStreamCreateWithFlags(&e->io, hipStreamNonBlocking) != hipSuccess) {
        destroy(e);
        PAMO_CHECK(false, "pamo serve/batch: stream creation failed");
    }
    for (Slot& s : e->slots) {
        int rc = pamo_ctx_create(&s.c2, device);
        if (rc == 0 && exact_guard) rc = pamo_ctx_set_option(s.c2, PAMO_OPT_EXACT_GUARD, exact_guard);
        if (rc == 0 && e->has_s3) rc = pamo_s3_create(&s.c3, device, &e->s3p);
        if (rc == 0 && hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess) rc = 1;
        if (rc != 0) {
            destroy(e);
            return rc;
        }
    }
    for (int i = 0; i < n_slots; ++i) e->slots[i].th = std::thread(worker, e, i);
    *out = e;
    return 0;
}

static long long submit(Engine* e, Job* j) {
    std::lock_guard<std::mutex> lk(e->mu);
    j->id = e->next_id++;
    j->t_submit = now_s();
    j->est = estimate_bytes(j->nv, j->nf, j->stage3);
    e->jobs[j->id] = j;
    e->queue.push_back(j);
    e->cv_work.notify_all();
    return j->id;
}

}  // namespace serve
}  // namespace pamo

using pamo::serve::Engine;
using pamo::serve::Job;

extern "C" {

// ------------------------------------------------------------------ batched form
int pamo_batch_run(int device, int n_slots, const pamo_s3_params* s3, int n_meshes, const float* d_verts,
                   const int* d_faces, const long long* vert_off, const long long* face_off, const double* ratio,
                   const int* use_stage3, int tolerance, float threshold, long long iters, long long min_verts,
                   int exact_guard, float* d_out_verts, int* d_out_faces, int* out_nv, int* out_nf,
                   double* seconds, pamo_batch_done_fn on_done, void* user) {
    pamo::clear_error();
    PAMO_CHECK(n_meshes >= 0 && d_verts && d_faces && vert_off && face_off && ratio && use_stage3 && d_out_verts &&
                   d_out_faces && out_nv && out_nf,
               "pamo_batch_run: null argument");
    for (int i = 0; i < n_meshes; ++i) {
        PAMO_CHECK(vert_off[i + 1] > vert_off[i] && face_off[i + 1] > face_off[i], "pamo_batch_run: empty mesh");
        PAMO_CHECK(!use_stage3[i] || s3, "pamo_batch_run: stage 3 requested without stage-3 parameters");
    }
    Engine* e = nullptr;
    PAMO_TRY(pamo::serve::create(&e, device, n_slots, 0, s3, exact_guard));
    e->on_done = on_done;
    e->on_done_user = user;
    const double t0 = pamo::serve::now_s();
    // longest first (LPT on face count)
    std::vector<int> order(n_meshes);
    for (int i = 0; i < n_meshes; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return face_off[a + 1] - face_off[a] > face_off[b + 1] - face_off[b];
    });
    std::vector<Job*> js(n_meshes);
    for (int k = 0; k < n_meshes; ++k) {
        const int i = order[k];
        Job* j = new Job();
        j->index = i;
        j->dv = d_verts + 3 * vert_off[i];
        j->df = d_faces + 3 * face_off[i];
        j->nv = (int)(vert_off[i + 1] - vert_off[i]);
        j->nf = (int)(face_off[i + 1] - face_off[i]);
        j->ratio = ratio[i];
        j->stage3 = use_stage3[i];
        j->tolerance = tolerance;
        j->threshold = threshold;
        j->iters = iters;
        j->min_verts = min_verts;
        j->out_v = d_out_verts + 3 * vert_off[i];
        j->out_f = d_out_faces + 3 * face_off[i];
        js[i] = j;
        pamo::serve::submit(e, j);
    }
    int rc = 0;
    std::string err;
    {
        std::unique_lock<std::mutex> lk(e->mu);
        e->cv_done.wait(lk, [&] {
            for (Job* j : js)
                if (!j->done) return false;
            return true;
        });
    }
    for (int i = 0; i < n_meshes; ++i) {
        out_nv[i] = js[i]->out_nv;
        out_nf[i] = js[i]->out_nf;
        if (js[i]->status != 0 && rc == 0) {
            rc = js[i]->status;
            err = "mesh " + std::to_string(i) + ": " + js[i]->err;
        }
    }
    if (seconds) *seconds = pamo::serve::now_s() - t0;
    pamo::serve::destroy(e);
    if (rc != 0) pamo::set_error("pamo_batch_run: " + err);
    return rc;
}

// ------------------------------------------------------------------ served form
int pamo_serve_create(pamo_server** out, int device, int n_slots, long long hbm_budget_bytes,
                      const pamo_s3_params* s3, int exact_guard) {
    pamo::clear_error();
    Engine* e = nullptr;
    PAMO_TRY(pamo::serve::create(&e, device, n_slots, hbm_budget_bytes, s3, exact_guard));
    *out = reinterpret_cast<pamo_server*>(e);
    return 0;
}

int pamo_serve_destroy(pamo_server* srv) {
    if (srv) pamo::serve::destroy(reinterpret_cast<Engine*>(srv));
    return 0;
}

int pamo_serve_submit(pamo_server* srv, const float* d_verts, int n_verts, const int* d_faces, int n_faces,
                      double ratio, int use_stage3, int tolerance, float threshold, long long iters,
                      long long min_verts, long long* job_id) {
    pamo::clear_error();
    PAMO_CHECK(srv && d_verts && d_faces && job_id && n_verts > 0 && n_faces > 0, "pamo_serve_submit: bad argument");
    Engine* e = reinterpret_cast<Engine*>(srv);
    PAMO_CHECK(!use_stage3 || e->has_s3, "pamo_serve_submit: server created without stage-3 parameters");
    Job* j = new Job();
    j->dv = d_verts;
    j->df = d_faces;
    j->nv = n_verts;
    j->nf = n_faces;
    j->ratio = ratio;
    j->stage3 = use_stage3;
    j->tolerance = tolerance;
    j->threshold = threshold;
    j->iters = iters;
    j->min_verts = min_verts;
    j->owns_out = true;
    hipError_t err = hipSetDevice(e->device);
    if (err == hipSuccess) err = hipMalloc((void**)&j->out_v, 12 * (size_t)n_verts);
    if (err == hipSuccess) err = hipMalloc((void**)&j->out_f, 12 * (size_t)n_faces);
    if (err != hipSuccess) {
        if (j->out_v) (void)hipFree(j->out_v);
        delete j;
        PAMO_HIP(err);
    }
    *job_id = pamo::serve::submit(e, j);
    return 0;
}

// state: 0 queued, 1 running, 2 done, 3 failed
int pamo_serve_poll(pamo_server* srv, long long job_id, int wait, int* state, int* n_verts_out, int* n_faces_out,
                    double* latency_s, double* queue_s) {
    pamo::clear_error();
    PAMO_CHECK(srv && state, "pamo_serve_poll: null argument");
    Engine* e = reinterpret_cast<Engine*>(srv);
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->jobs.find(job_id);
    PAMO_CHECK(it != e->jobs.end(), "pamo_serve_poll: unknown job");
    Job* j = it->second;
    if (wait) e->cv_done.wait(lk, [&] { return j->done; });
    *state = j->done ? (j->status ? 3 : 2) : (j->t_start > 0 ? 1 : 0);
    if (n_verts_out) *n_verts_out = j->out_nv;
    if (n_faces_out) *n_faces_out = j->out_nf;
    if (latency_s) *latency_s = j->done ? j->t_end - j->t_submit : 0.0;
    if (queue_s) *queue_s = j->t_start > 0 ? j->t_start - j->t_submit : 0.0;
    if (j->done && j->status) pamo::set_error(j->err);
    return 0;
}

// Copies a finished job's result (device pointers, sized by poll's counts)
// and releases the job.
int pamo_serve_fetch(pamo_server* srv, long long job_id, float* d_out_verts, int* d_out_faces) {
    pamo::clear_error();
    PAMO_CHECK(srv && d_out_verts && d_out_faces, "pamo_serve_fetch: null argument");
    Engine* e = reinterpret_cast<Engine*>(srv);
    Job* j = nullptr;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        auto it = e->jobs.find(job_id);
        PAMO_CHECK(it != e->jobs.end(), "pamo_serve_fetch: unknown job");
        j = it->second;
        PAMO_CHECK(j->done, "pamo_serve_fetch: job not finished");
        e->jobs.erase(it);
    }
    PAMO_HIP(hipSetDevice(e->device));
    hipError_t er = hipMemcpyAsync(d_out_verts, j->out_v, 12 * (size_t)j->out_nv, hipMemcpyDeviceToDevice, e->io);
    if (er == hipSuccess)
        er = hipMemcpyAsync(d_out_faces, j->out_f, 12 * (size_t)j->out_nf, hipMemcpyDeviceToDevice, e->io);
    if (er == hipSuccess) er = hipStreamSynchronize(e->io);
    const int status = j->status;
    const std::string err = j->err;
    (void)hipFree(j->out_v);
    (void)hipFree(j->out_f);
    delete j;
    PAMO_HIP(er);
    if (status) {
        pamo::set_error("pamo_serve_fetch: job failed: " + err);
        return status;
    }
    return 0;
}

}  // extern "C"
