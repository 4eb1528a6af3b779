// Request server: continuous batching of per-CTU motion-estimation requests
// from several concurrent streams (no counterpart in the reference, whose
// encoder runs ME per PU on its own threads: inter.c:834-970).
//
// Threads:
//   callers     put_picture / submit / wait / release_picture (any thread)
//   dispatcher  forms a batch from the pending queue whenever a batch slot is
//               free, uploads its job list and picture descriptors, enqueues
//               the search kernels and the result copies on the slot's stream
//   completer   waits for batches in launch order, scatters their results to
//               the callers' buffers, retires tickets and pictures
//
// A batch may hold CTUs of any stream and picture: the kernels take a job
// list (picture index << 16 | CTU) and a picture descriptor per distinct
// (current, references) combination, so a CTU's result does not depend on
// the rest of its batch.  Pictures live in device memory (me_ref_s) and are
// admitted against the configured HBM budget; released pictures go to a
// pool and are reused by pictures of the same size without hipMalloc/hipFree.
#include "frame.h"

#include <string.h>
#include <time.h>
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <set>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

ME_EXPORT double me_server_now_us(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e6 + (double)ts.tv_nsec * 1e-3;
}

namespace {

struct Picture {
    me_ref_s *ref = nullptr;
    int stream = 0;
    int64_t poc = 0;
    int pending = 0;        // submitted requests naming it that have not completed
    bool released = false;
};

struct Stream {
    int w, h, ctu_cols, n_ctu;
    std::map<int64_t, Picture *> pics;
    std::set<int64_t> loading;   // POCs reserved by a put that is still uploading
};

struct Job {
    int64_t ticket;
    int stream;
    int ctu;
    Picture *cur;
    Picture *refs[ME_MAX_REFS];
    me_pu_result_t *res;
    int8_t *best;
    double submit_us;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    u32 *d_jobs = nullptr, *h_jobs = nullptr;
    PicDesc *d_pics = nullptr, *h_pics = nullptr;
    u64 *d_keys = nullptr;
    me_pu_result_t *d_res = nullptr, *h_res = nullptr;
    s8 *d_best = nullptr, *h_best = nullptr;
    void *hv_res = nullptr, *hv_best = nullptr;   // device views of h_res, h_best
    std::vector<Job> jobs;
    int rc = 0;
    bool busy = false;
    double launch_us = 0;   // when the dispatcher began enqueuing the batch
};

struct Ticket {
    bool done = false;
    int rc = 0;
    double done_us = 0;
};

size_t picture_bytes(int w, int h)
{
    const size_t planes = 16 * (size_t)(w + 2 * ME_PAD) * (h + 2 * ME_PAD) * sizeof(pel);
    const size_t luma = (size_t)w * h * sizeof(pel) + 64;
    const size_t tmp = 3 * (size_t)(w + 16) * (h + 16) * sizeof(s16);
    return planes + luma + tmp;
}

}  // namespace

struct me_server_s {
    me_server_cfg_t cfg;
    MeSearch search;
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work;   // dispatcher: new work or a free slot
    std::condition_variable cv_done;   // completer, waiters, admission
    std::vector<Stream> streams;
    std::deque<Job> pending;
    std::vector<Slot> slots;
    std::deque<int> inflight;          // slot indices in launch order
    std::unordered_map<int64_t, Ticket> tickets;
    int64_t next_ticket = 1;
    uint64_t used_bytes = 0;
    std::vector<me_ref_s *> pool;      // resident buffers of released pictures
    hipStream_t upload = nullptr;
    std::thread dispatcher, completer;
    bool stop = false, dispatch_exited = false;
    me_server_stats_t stats;
};

static void slot_free(Slot &s)
{
    if (s.st) (void)hipStreamDestroy(s.st);
    if (s.done) (void)hipEventDestroy(s.done);
    (void)hipFree(s.d_jobs);
    (void)hipFree(s.d_pics);
    (void)hipFree(s.d_keys);
    (void)hipFree(s.d_res);
    (void)hipFree(s.d_best);
    (void)hipHostFree(s.h_jobs);
    (void)hipHostFree(s.h_pics);
    (void)hipHostFree(s.h_res);
    (void)hipHostFree(s.h_best);
}

static int slot_alloc(Slot &s, int max_jobs, int npu, int nref)
{
    const size_t nres = (size_t)max_jobs * npu * nref;
    ME_CHECK(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
    ME_CHECK(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    ME_CHECK(hipMalloc((void **)&s.d_jobs, max_jobs * sizeof(u32)));
    ME_CHECK(hipMalloc((void **)&s.d_pics, max_jobs * sizeof(PicDesc)));
    ME_CHECK(hipMalloc((void **)&s.d_keys, nres * sizeof(u64)));
    ME_CHECK(hipMalloc((void **)&s.d_res, nres * sizeof(me_pu_result_t)));
    ME_CHECK(hipMalloc((void **)&s.d_best, (size_t)max_jobs * npu));
    ME_CHECK(hipHostMalloc((void **)&s.h_jobs, max_jobs * sizeof(u32), hipHostMallocDefault));
    ME_CHECK(hipHostMalloc((void **)&s.h_pics, max_jobs * sizeof(PicDesc), hipHostMallocDefault));
    ME_CHECK(hipHostMalloc((void **)&s.h_res, nres * sizeof(me_pu_result_t), hipHostMallocDefault));
    ME_CHECK(hipHostMalloc((void **)&s.h_best, (size_t)max_jobs * npu, hipHostMallocDefault));
    ME_CHECK(hipHostGetDevicePointer(&s.hv_res, s.h_res, 0));
    ME_CHECK(hipHostGetDevicePointer(&s.hv_best, s.h_best, 0));
    return 0;
}

// Enqueues one batch on its slot (called by the dispatcher without the lock).
// A batch's results to the slot's pinned host buffers, written by the GPU
// itself: a runtime device-to-host copy above 16 KB now and then took
// milliseconds to enqueue, which every request of the batch then waited for.
__global__ void k_copy_out(uint4 *__restrict__ hres, const uint4 *__restrict__ dres, size_t n16,
                           u8 *__restrict__ hbest, const u8 *__restrict__ dbest, size_t nb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x, step = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n16; k += step) hres[k] = dres[k];
    for (size_t k = i; k < nb; k += step) hbest[k] = dbest[k];
}

static void copy_out(Slot &sl, size_t nres, size_t nbest)
{
    const size_t n16 = nres * sizeof(me_pu_result_t) / 16;
    const size_t blocks = std::min<size_t>(64, (std::max(n16, nbest) + 255) / 256);
    hipLaunchKernelGGL(k_copy_out, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, sl.st, (uint4 *)sl.hv_res,
                       (const uint4 *)sl.d_res, n16, (u8 *)sl.hv_best, (const u8 *)sl.d_best, nbest);
}

static int launch_batch(me_server_s *S, Slot &sl)
{
    const int n = (int)sl.jobs.size();
    const int nref = S->search.nref;
    // one descriptor per distinct (current picture, references)
    std::map<std::vector<const void *>, int> index;
    int npics = 0;
    for (int j = 0; j < n; j++) {
        const Job &jb = sl.jobs[j];
        std::vector<const void *> key(1 + nref);
        key[0] = jb.cur;
        for (int r = 0; r < nref; r++) key[1 + r] = jb.refs[r];
        auto it = index.find(key);
        int pi;
        if (it == index.end()) {
            pi = npics++;
            index.emplace(key, pi);
            PicDesc &d = sl.h_pics[pi];
            memset(&d, 0, sizeof(d));
            const Stream &st = S->streams[jb.stream];
            d.cur = jb.cur->ref->luma;
            for (int r = 0; r < nref; r++) d.planes[r] = jb.refs[r]->ref->planes;
            d.plane_elems = jb.refs[0]->ref->plane_elems;
            d.ref_stride = jb.refs[0]->ref->stride;
            d.pic_w = st.w;
            d.pic_h = st.h;
            d.ctu_cols = st.ctu_cols;
        } else {
            pi = it->second;
        }
        sl.h_jobs[j] = ((u32)pi << 16) | (u32)jb.ctu;
    }
    ME_CHECK(hipMemcpyAsync(sl.d_jobs, sl.h_jobs, n * sizeof(u32), hipMemcpyHostToDevice, sl.st));
    ME_CHECK(hipMemcpyAsync(sl.d_pics, sl.h_pics, npics * sizeof(PicDesc), hipMemcpyHostToDevice, sl.st));
    JobSet js;
    js.pics = sl.d_pics;
    js.jobs = sl.d_jobs;
    js.n_jobs = n;
    // batches are a few CTUs: one workgroup per band, keys merged with atomics
    KeyLayout kl;
    me_search_keys(&S->search, n, 1, &kl);
    if (me_search_enqueue(&S->search, js, kl, sl.d_keys, sl.d_res, sl.d_best, sl.st, nullptr)) return -1;
    const size_t nres = (size_t)n * S->search.npu * nref;
    copy_out(sl, nres, (size_t)n * S->search.npu);
    ME_CHECK(hipGetLastError());
    ME_CHECK(hipEventRecord(sl.done, sl.st));
    return 0;
}

static void dispatch_loop(me_server_s *S)
{
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        int slot = -1;
        for (;;) {
            if (S->pending.empty()) {
                if (S->stop) {
                    S->dispatch_exited = true;
                    S->cv_done.notify_all();
                    return;
                }
                S->cv_work.wait(lk);
                continue;
            }
            slot = -1;
            for (size_t i = 0; i < S->slots.size(); i++)
                if (!S->slots[i].busy) { slot = (int)i; break; }
            if (slot < 0) { S->cv_work.wait(lk); continue; }
            // launch at once when the device is idle, the batch is full, or
            // the oldest request has waited max_wait_us; else let it fill
            const double age = me_server_now_us() - S->pending.front().submit_us;
            if (S->inflight.empty() || (int)S->pending.size() >= S->cfg.max_batch_ctus || S->stop ||
                age >= S->cfg.max_wait_us)
                break;
            S->cv_work.wait_for(lk, std::chrono::microseconds((int64_t)(S->cfg.max_wait_us - age) + 1));
        }
        Slot &sl = S->slots[slot];
        sl.busy = true;
        sl.jobs.clear();
        const int n = std::min((int)S->pending.size(), S->cfg.max_batch_ctus);
        bool mixed = false;
        for (int j = 0; j < n; j++) {
            sl.jobs.push_back(S->pending.front());
            S->pending.pop_front();
            mixed |= sl.jobs[j].stream != sl.jobs[0].stream;
        }
        lk.unlock();
        sl.launch_us = me_server_now_us();
        sl.rc = launch_batch(S, sl);
        const double enq_us = me_server_now_us() - sl.launch_us;
        lk.lock();
        S->stats.max_launch_us = std::max<uint64_t>(S->stats.max_launch_us, (uint64_t)enq_us);
        S->inflight.push_back(slot);
        S->stats.batches++;
        S->stats.mixed_batches += mixed ? 1 : 0;
        S->stats.max_batch_seen = std::max<uint64_t>(S->stats.max_batch_seen, (uint64_t)n);
        S->stats.max_inflight_seen = std::max<uint64_t>(S->stats.max_inflight_seen, (uint64_t)S->inflight.size());
        S->cv_done.notify_all();
    }
}

// Drops a picture that is released and no longer named by a pending
// request: its buffers go to the pool (the lock is held).
static void picture_retire_if_done(me_server_s *S, Picture *p)
{
    if (!p->released || p->pending > 0) return;
    Stream &st = S->streams[p->stream];
    st.pics.erase(p->poc);
    S->pool.push_back(p->ref);
    delete p;
}

static void complete_loop(me_server_s *S)
{
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        while (S->inflight.empty()) {
            if (S->dispatch_exited) return;
            S->cv_done.wait(lk);
        }
        const int slot = S->inflight.front();
        Slot &sl = S->slots[slot];
        lk.unlock();
        int rc = sl.rc;
        if (rc == 0) {
            hipError_t e = hipEventSynchronize(sl.done);
            if (e != hipSuccess) { me_set_error("server batch", e, __FILE__, __LINE__); rc = -1; }
        } else {
            (void)hipStreamSynchronize(sl.st);   // whatever part of the batch was enqueued
        }
        const int npu = S->search.npu, nref = S->search.nref;
        if (rc == 0) {
            for (size_t j = 0; j < sl.jobs.size(); j++) {
                const Job &jb = sl.jobs[j];
                if (jb.res) memcpy(jb.res, sl.h_res + j * npu * nref, (size_t)npu * nref * sizeof(me_pu_result_t));
                if (jb.best) memcpy(jb.best, sl.h_best + j * npu, (size_t)npu);
            }
        }
        const double now = me_server_now_us();
        lk.lock();
        S->stats.max_batch_us = std::max<uint64_t>(S->stats.max_batch_us, (uint64_t)(now - sl.launch_us));
        S->inflight.pop_front();
        for (const Job &jb : sl.jobs) {
            auto it = S->tickets.find(jb.ticket);
            if (it != S->tickets.end()) {
                it->second.done = true;
                it->second.rc = rc;
                it->second.done_us = now;
            }
            jb.cur->pending--;
            picture_retire_if_done(S, jb.cur);
            for (int r = 0; r < nref; r++) {
                jb.refs[r]->pending--;
                picture_retire_if_done(S, jb.refs[r]);
            }
            S->stats.requests_done++;
        }
        sl.jobs.clear();
        sl.busy = false;
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

ME_EXPORT int me_server_create(me_server_t *out, const me_server_cfg_t *cfg)
{
    if (me_hip_init(-1)) return -1;
    if (!out || !cfg || cfg->max_batch_ctus < 1 || cfg->max_batch_ctus > 0xFFFF || cfg->max_inflight < 1 ||
        cfg->max_inflight > 8 || cfg->max_wait_us < 0) {
        me_set_error("me_server_create: bad configuration", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    me_server_s *S = new me_server_s();
    S->cfg = *cfg;
    S->device = me_hip_device();
    memset(&S->stats, 0, sizeof(S->stats));
    // every stream must satisfy R <= max(W,H) - 4, so the window's uniform
    // part is [-R, R] for all of them (checked at me_server_open_stream).
    // Batches are a few CTUs: sub-pel positions run 4 at a time for latency.
    if (me_search_init(&S->search, cfg->search_range, cfg->lambda_mv, cfg->shape_mask, cfg->num_refs, 0,
                       cfg->search_range + 4, cfg->band_rows, cfg->waves, cfg->dx_cols, 4)) {
        delete S;
        return -1;
    }
    S->slots.resize(cfg->max_inflight);
    int rc = 0;
    for (auto &sl : S->slots)
        if (!rc) rc = slot_alloc(sl, cfg->max_batch_ctus, S->search.npu, S->search.nref);
    if (!rc) {
        hipError_t e = hipStreamCreateWithFlags(&S->upload, hipStreamNonBlocking);
        if (e != hipSuccess) { me_set_error("me_server_create", e, __FILE__, __LINE__); rc = -1; }
    }
    // the first requests should not pay for kernel loading, queue creation
    // or the first copies: a batch's copies, once, on every slot (the
    // buffers' contents do not matter yet)
    for (auto &sl : S->slots) {
        if (!rc) rc = me_search_warmup(&S->search, sl.st);
        if (rc) break;
        // whole buffers: copies of different sizes take different paths
        const size_t mj = (size_t)cfg->max_batch_ctus, np = S->search.npu, nres = mj * np * S->search.nref;
        hipError_t e = hipMemcpyAsync(sl.d_jobs, sl.h_jobs, mj * sizeof(u32), hipMemcpyHostToDevice, sl.st);
        if (e == hipSuccess) e = hipMemcpyAsync(sl.d_pics, sl.h_pics, mj * sizeof(PicDesc), hipMemcpyHostToDevice, sl.st);
        if (e == hipSuccess) {
            copy_out(sl, nres, mj * np);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(sl.done, sl.st);
        if (e == hipSuccess) e = hipEventSynchronize(sl.done);
        if (e != hipSuccess) { me_set_error("me_server_create: warm-up", e, __FILE__, __LINE__); rc = -1; }
    }
    if (!rc) rc = me_search_warmup(&S->search, S->upload);
    if (rc) {
        for (auto &sl : S->slots) slot_free(sl);
        me_search_free(&S->search);
        delete S;
        return -1;
    }
    S->dispatcher = std::thread(dispatch_loop, S);
    S->completer = std::thread(complete_loop, S);
    *out = S;
    return 0;
}

ME_EXPORT int me_server_open_stream(me_server_t S, int pic_w, int pic_h, int *stream_id)
{
    if (!S || !stream_id || pic_w < 1 || pic_h < 1 || pic_w > 128 * 0xFFFF ||
        std::max(pic_w, pic_h) - 4 < S->cfg.search_range || ((pic_w + 127) / 128) * ((pic_h + 127) / 128) > 0xFFFF) {
        me_set_error("me_server_open_stream: bad picture size (needs max(W,H) - 4 >= search range)",
                     hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    Stream st;
    st.w = pic_w;
    st.h = pic_h;
    st.ctu_cols = (pic_w + 127) / 128;
    st.n_ctu = st.ctu_cols * ((pic_h + 127) / 128);
    S->streams.push_back(st);
    *stream_id = (int)S->streams.size() - 1;
    return 0;
}

static int put_picture_locked(me_server_s *S, int sid, int64_t poc, const me_pel *luma, int stride, int timeout_ms,
                              std::unique_lock<std::mutex> &lk);

ME_EXPORT int me_server_put_picture(me_server_t S, int sid, int64_t poc, const me_pel *luma, int stride, int timeout_ms)
{
    if (!S || !luma) return -1;
    std::unique_lock<std::mutex> lk(S->mu);
    if (sid < 0 || sid >= (int)S->streams.size() || stride < S->streams[sid].w || S->streams[sid].pics.count(poc) ||
        S->streams[sid].loading.count(poc)) {
        me_set_error("me_server_put_picture: bad stream, stride or duplicate picture", hipErrorInvalidValue, __FILE__,
                     __LINE__);
        return -1;
    }
    // The POC stays reserved while the lock is dropped (allocation, upload),
    // so a concurrent put of the same picture fails instead of replacing it.
    S->streams[sid].loading.insert(poc);
    const int rc = put_picture_locked(S, sid, poc, luma, stride, timeout_ms, lk);
    S->streams[sid].loading.erase(poc);
    return rc;
}

// The body of me_server_put_picture with the POC reserved; lk is held on
// entry and on return.
static int put_picture_locked(me_server_s *S, int sid, int64_t poc, const me_pel *luma, int stride, int timeout_ms,
                              std::unique_lock<std::mutex> &lk)
{
    const int w = S->streams[sid].w, h = S->streams[sid].h;
    const size_t need = picture_bytes(w, h);
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms);
    me_ref_s *ref = nullptr;
    bool waited = false;
    for (;;) {
        auto it = std::find_if(S->pool.begin(), S->pool.end(), [&](me_ref_s *r) { return r->w == w && r->h == h; });
        if (it != S->pool.end()) {
            ref = *it;
            S->pool.erase(it);
            break;
        }
        if (S->used_bytes + need <= S->cfg.hbm_budget) {
            S->used_bytes += need;
            lk.unlock();
            int rc = me_ref_create(&ref, w, h);
            lk.lock();
            if (rc) {
                S->used_bytes -= need;
                return -1;
            }
            break;
        }
        if (!S->pool.empty()) {   // free a pooled picture of another size
            me_ref_s *r = S->pool.back();
            S->pool.pop_back();
            S->used_bytes -= picture_bytes(r->w, r->h);
            me_ref_destroy(r);
            continue;
        }
        if (!waited) { S->stats.admission_waits++; waited = true; }
        if (timeout_ms < 0) S->cv_done.wait(lk);
        else if (S->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) return 1;
    }
    lk.unlock();
    int rc = me_ref_load(ref, luma, stride, 0, S->upload);
    lk.lock();
    if (rc) {
        S->pool.push_back(ref);
        return -1;
    }
    Picture *p = new Picture();
    p->ref = ref;
    p->stream = sid;
    p->poc = poc;
    S->streams[sid].pics[poc] = p;
    return 0;
}

ME_EXPORT int me_server_release_picture(me_server_t S, int sid, int64_t poc)
{
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    if (sid < 0 || sid >= (int)S->streams.size()) return -1;
    auto it = S->streams[sid].pics.find(poc);
    if (it == S->streams[sid].pics.end() || it->second->released) {
        me_set_error("me_server_release_picture: no such picture", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    it->second->released = true;
    picture_retire_if_done(S, it->second);
    S->cv_done.notify_all();
    return 0;
}

ME_EXPORT int64_t me_server_submit(me_server_t S, int sid, int64_t cur_poc, int ctu, const int64_t *ref_pocs,
                                   me_pu_result_t *res, int8_t *best_ref)
{
    if (!S || !ref_pocs) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    if (sid < 0 || sid >= (int)S->streams.size() || ctu < 0 || ctu >= S->streams[sid].n_ctu) {
        me_set_error("me_server_submit: bad stream or CTU", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    Stream &st = S->streams[sid];
    Job jb;
    memset(&jb, 0, sizeof(jb));
    auto find = [&](int64_t poc) -> Picture * {
        auto it = st.pics.find(poc);
        return it == st.pics.end() || it->second->released ? nullptr : it->second;
    };
    jb.cur = find(cur_poc);
    bool ok = jb.cur != nullptr;
    for (int r = 0; r < S->search.nref && ok; r++) ok = (jb.refs[r] = find(ref_pocs[r])) != nullptr;
    if (!ok) {
        me_set_error("me_server_submit: picture not resident (or released)", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    jb.ticket = S->next_ticket++;
    jb.stream = sid;
    jb.ctu = ctu;
    jb.res = res;
    jb.best = best_ref;
    jb.submit_us = me_server_now_us();
    jb.cur->pending++;
    for (int r = 0; r < S->search.nref; r++) jb.refs[r]->pending++;
    S->tickets.emplace(jb.ticket, Ticket());
    S->pending.push_back(jb);
    S->cv_work.notify_one();
    return jb.ticket;
}

ME_EXPORT int me_server_wait(me_server_t S, int64_t ticket, int timeout_ms, double *done_us)
{
    if (!S) return -1;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) {
        me_set_error("me_server_wait: unknown ticket", hipErrorInvalidValue, __FILE__, __LINE__);
        return -1;
    }
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms);
    while (!it->second.done) {
        if (timeout_ms < 0) S->cv_done.wait(lk);
        else if (S->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) return 1;
        it = S->tickets.find(ticket);
    }
    const int rc = it->second.rc;
    if (done_us) *done_us = it->second.done_us;
    S->tickets.erase(it);
    return rc ? -1 : 0;
}

ME_EXPORT int me_server_stats(me_server_t S, me_server_stats_t *out)
{
    if (!S || !out) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    out->picture_bytes = S->used_bytes;
    return 0;
}

ME_EXPORT void me_server_destroy(me_server_t S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
    if (S->dispatcher.joinable()) S->dispatcher.join();
    if (S->completer.joinable()) S->completer.join();
    for (auto &sl : S->slots) slot_free(sl);
    for (auto &st : S->streams)
        for (auto &kv : st.pics) {
            me_ref_destroy(kv.second->ref);
            delete kv.second;
        }
    for (me_ref_s *r : S->pool) me_ref_destroy(r);
    if (S->upload) (void)hipStreamDestroy(S->upload);
    me_search_free(&S->search);
    delete S;
}
