// bmi_server.hip -- request server: continuous batching of decode requests
// over a fixed HBM budget.
//
// One scheduler thread owns the device side.  It admits waiting requests
// in arrival order while their footprint fits the free part of the arena
// (one hipMalloc made at create), decodes each admitted set as one batch
// plan on a free stream (copy-in, plan, copy-out), and polls the batches
// in flight; a request that arrives while batches run joins the next batch
// as soon as a stream and memory are free, so its work overlaps theirs on
// the device.  Callers submit from any thread and wait on a ticket.
//
// The server has no counterpart in the reference, whose single CUDA host
// thread decodes one tile job at a time (cudainterface.cpp:2276-2645).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "bmi_arena.h"
#include "bmi_internal.h"

namespace {

using bmi::Arena;

constexpr size_t kAlign = 256;
inline size_t al(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

int pixel_bytes(int fmt)
{
    switch (fmt) {
    case IMAGE_FOTMAT_32BITS_ARGB:
    case IMAGE_FOTMAT_32BITS_4COMPONENTS: return 4;
    case IMAGE_FOTMAT_48BITS_RGB:
    case IMAGE_FOTMAT_48BITS_3COMPONENTS: return 6;
    case IMAGE_FOTMAT_24BITS_RGB:
    case IMAGE_FOTMAT_24BITS_3COMPONENTS: return 3;
    case IMAGE_FOTMAT_64BITS_4COMPONENTS:
    case IMAGE_FOTMAT_64BITS_ARGB: return 8;
    default: return 0;
    }
}

int ceil_shift_h(int v, int s) { return (v + (1 << s) - 1) >> s; }

// size of the decoded image (reduced when top_level < levels)
void output_size(const BmiTileDesc &t, int *ow, int *oh)
{
    if (t.top_level < 0 || t.top_level >= t.levels) {
        *ow = t.width;
        *oh = t.height;
        return;
    }
    const int s = t.levels - t.top_level;
    *ow = ceil_shift_h(t.x0 + t.width, s) - ceil_shift_h(t.x0, s);
    *oh = ceil_shift_h(t.y0 + t.height, s) - ceil_shift_h(t.y0, s);
}

struct Req {
    unsigned long long ticket = 0;
    BmiRequest r{};
    size_t io_bytes = 0;    // coefficient + output copies (host_io) or plain-path copies
    size_t plan_bytes = 0;  // single-request plan estimate
    int dev_pitch = 0;      // elements, device coefficient planes
    int ow = 0, oh = 0;
    BmiRequestResult res{};
    bool done = false;
};

struct Batch {
    int id = 0;
    int stream = 0;
    std::vector<Req *> reqs;
    BmiBatchPlan *plan = nullptr;
    size_t off = Arena::kNone;
    hipEvent_t ev = nullptr;
};

double now_us(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// initial pinned staging per stream: the job and item tables of a batch of
// 32 requests of up to 8192x8192 pixels
constexpr size_t kStagingBytes = size_t(4) << 20;

struct BmiServer {
    BmiServerConfig cfg{};
    int device = 0;
    void *base = nullptr;
    Arena arena;
    std::vector<hipStream_t> streams;
    std::vector<bool> stream_busy;
    std::vector<void *> staging;        // per stream: pinned copy of plan tables
    std::vector<size_t> staging_bytes;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work;  // scheduler: new request or shutdown
    std::condition_variable cv_done;  // waiters: a request completed
    std::deque<Req *> pending;
    std::map<unsigned long long, Req *> reqs;
    unsigned long long next_ticket = 1;
    bool stop = false;
    BmiServerStats stats{};
    int next_batch = 1;

    std::vector<Batch> inflight;  // scheduler thread only
    std::thread thread;

    void run();
    bool launch(std::vector<Req *> &set);
    void complete(Batch &b, int status);
};

// ---------------------------------------------------------------------------

bool BmiServer::launch(std::vector<Req *> &set)
{
    // find a free stream
    int si = -1;
    for (int i = 0; i < (int)streams.size(); ++i)
        if (!stream_busy[i]) {
            si = i;
            break;
        }
    if (si < 0) return false;
    hipStream_t s = streams[si];
    const bool plain = cfg.mode == BMI_SERVE_PLAIN;

    // The batch's block holds the requests' copies first, then the plan.
    // Its size is the sum of the admitted footprints (plan part: the
    // per-request upper bounds), so the plan built below always fits.
    size_t io = 0, plan_res = 0;
    for (Req *q : set) {
        io += q->io_bytes;
        plan_res += q->plan_bytes;
    }
    const size_t off = arena.alloc(io + plan_res, kAlign);
    if (off == Arena::kNone) return false;
    Batch b;
    b.id = next_batch++;
    b.stream = si;
    b.reqs = set;
    b.off = off;
    char *blk = (char *)base + off;
    const double t_start = now_us(t0);
    int rc = BMI_OK;

    // carve per-request copies and rewrite the descriptors to device buffers
    std::vector<BmiTileDesc> d(set.size());
    char *p = blk;
    std::vector<void *> temp(set.size(), nullptr);
    for (size_t i = 0; i < set.size() && !rc; ++i) {
        Req &q = *set[i];
        BmiTileDesc &t = d[i];
        t = q.r.tile;
        const BmiTileDesc &src = q.r.tile;
        const size_t plane = (size_t)q.dev_pitch * src.height * 4;
        if (q.r.host_io || plain) {
            const hipMemcpyKind k = q.r.host_io ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
            for (int c = 0; c < src.ncomp && !rc; ++c) {
                void *dst = p + (size_t)c * al(plane);
                if (hipMemcpy2DAsync(dst, (size_t)q.dev_pitch * 4, src.coef[c],
                                     (size_t)src.coef_pitch * 4, (size_t)src.width * 4, src.height,
                                     k, s) != hipSuccess)
                    rc = BMI_E_HIP;
                t.coef[c] = dst;
            }
            t.coef_pitch = q.dev_pitch;
            p += (size_t)src.ncomp * al(plane);
        }
        if (q.r.host_io) {
            t.out = p;
            t.out_pitch_px = q.ow;
            t.out_x = t.out_y = 0;
            p += al((size_t)q.ow * q.oh * pixel_bytes(src.out_format));
        }
        if (plain) {
            temp[i] = p;
            p += al((size_t)src.width * src.height * 4 + 256);
        }
    }
    if (!rc) {
        if (plain) {
            for (size_t i = 0; i < set.size() && !rc; ++i)
                rc = BMI_plain_decode(&d[i], temp[i],
                                      (size_t)set[i]->r.tile.width * set[i]->r.tile.height * 4 + 256,
                                      s);
        } else {
            BmiBatchPlan *plan = nullptr;
            rc = bmi_plan_build(&plan, d.data(), (int)d.size(), &cfg.tuning);
            b.plan = plan;
            if (!rc && bmi_plan_device_bytes(plan) > plan_res) rc = BMI_E_NOMEM;
            if (!rc) {
                // pinned staging of this stream (free: the stream is idle)
                const size_t tb = bmi_plan_table_bytes(plan);
                if (staging_bytes[si] < tb) {
                    if (staging[si]) (void)hipHostFree(staging[si]);
                    staging[si] = nullptr;
                    staging_bytes[si] = 0;
                    const size_t want = std::max(tb + tb / 2, (size_t)1 << 20);
                    if (hipHostMalloc(&staging[si], want, 0) == hipSuccess) staging_bytes[si] = want;
                }
                rc = bmi_plan_bind(plan, p, s, staging[si], staging_bytes[si]);
            }
            if (!rc) rc = BMI_batch_plan_execute(plan, s);
        }
    }
    // copy results out
    for (size_t i = 0; i < set.size() && !rc; ++i) {
        Req &q = *set[i];
        if (!q.r.host_io) continue;
        const BmiTileDesc &src = q.r.tile;
        const int pb = pixel_bytes(src.out_format);
        char *host = (char *)src.out + ((size_t)src.out_y * src.out_pitch_px + src.out_x) * pb;
        if (hipMemcpy2DAsync(host, (size_t)src.out_pitch_px * pb, d[i].out, (size_t)q.ow * pb,
                             (size_t)q.ow * pb, q.oh, hipMemcpyDeviceToHost, s) != hipSuccess)
            rc = BMI_E_HIP;
    }
    if (!rc && (hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(b.ev, s) != hipSuccess))
        rc = BMI_E_HIP;
    for (Req *q : set) {
        q->res.start_us = t_start;
        q->res.batch_id = b.id;
        q->res.batch_size = (int)set.size();
    }
    if (rc) {
        (void)hipStreamSynchronize(s);
        complete(b, rc);
        return true;
    }
    stream_busy[si] = true;
    inflight.push_back(b);
    {
        std::lock_guard<std::mutex> g(mu);
        stats.batches++;
        stats.peak_bytes = std::max(stats.peak_bytes, arena.capacity() - arena.free_bytes());
        stats.peak_inflight = std::max(stats.peak_inflight, (int)inflight.size());
    }
    return true;
}

// release the batch's memory and publish its requests' results
void BmiServer::complete(Batch &b, int status)
{
    if (b.plan) BMI_batch_plan_destroy(b.plan);
    b.plan = nullptr;
    if (b.off != Arena::kNone) arena.release(b.off);
    if (b.ev) (void)hipEventDestroy(b.ev);
    b.ev = nullptr;
    const double t = now_us(t0);
    std::lock_guard<std::mutex> g(mu);
    for (Req *q : b.reqs) {
        q->res.status = status;
        q->res.done_us = t;
        if (!q->res.batch_id) q->res.batch_id = b.id;
        q->done = true;
        if (status == BMI_OK)
            stats.completed++;
        else
            stats.failed++;
    }
    cv_done.notify_all();
}

void BmiServer::run()
{
    (void)hipSetDevice(device);
    std::vector<size_t> foot;
    for (;;) {
        // 1. retire finished batches
        for (size_t i = 0; i < inflight.size();) {
            hipError_t e = hipEventQuery(inflight[i].ev);
            if (e == hipErrorNotReady) {
                ++i;
                continue;
            }
            if (cfg.tuning.persistent == 1 && e == hipSuccess && inflight[i].plan &&
                BMI_batch_plan_error_flag(inflight[i].plan) != 0)
                e = hipErrorLaunchFailure;
            stream_busy[inflight[i].stream] = false;
            complete(inflight[i], e == hipSuccess ? BMI_OK : BMI_E_HIP);
            inflight.erase(inflight.begin() + i);
        }
        // 2. admit the longest fitting prefix of the waiting requests
        std::vector<Req *> set;
        bool stopping;
        {
            std::unique_lock<std::mutex> g(mu);
            if (pending.empty() && inflight.empty()) {
                if (stop) return;
                cv_work.wait(g, [&] { return stop || !pending.empty(); });
                if (pending.empty() && stop) return;
            }
            stopping = stop;
            bool free_stream = false;
            for (bool busy : stream_busy) free_stream = free_stream || !busy;
            if (free_stream && !pending.empty()) {
                foot.clear();
                for (Req *q : pending) foot.push_back(q->plan_bytes + q->io_bytes);
                const int n = bmi::admit_prefix(foot, arena.largest_free(),
                                                std::max(1, cfg.max_batch));
                for (int i = 0; i < n; ++i) {
                    set.push_back(pending.front());
                    pending.pop_front();
                }
            }
        }
        (void)stopping;
        if (!set.empty()) {
            // the batch plan can need a little more than the sum of the
            // single-request estimates (alignment); shrink until it fits
            while (!set.empty() && !launch(set)) {
                std::lock_guard<std::mutex> g(mu);
                pending.push_front(set.back());
                set.pop_back();
            }
            continue;
        }
        if (!inflight.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

// ---------------------------------------------------------------------------

extern "C" int BMI_server_create(BmiServer **out, const BmiServerConfig *cfg)
{
    if (!out || !cfg || cfg->hbm_budget < (1u << 20) || cfg->num_streams < 1 ||
        cfg->num_streams > 8 || cfg->max_batch < 1 ||
        (cfg->mode != BMI_SERVE_FUSED && cfg->mode != BMI_SERVE_PLAIN))
        return bmi_fail(BMI_E_INVALID, "server_create: bad config");
    *out = nullptr;
    BmiServer *S = new BmiServer();
    S->cfg = *cfg;
    if (S->cfg.mode == BMI_SERVE_PLAIN) {
        S->cfg.num_streams = 1;
        S->cfg.max_batch = 1;
    }
    if (hipGetDevice(&S->device) != hipSuccess) {
        delete S;
        return bmi_fail(BMI_E_NODEVICE, "server_create: no device");
    }
    if (hipMalloc(&S->base, S->cfg.hbm_budget) != hipSuccess) {
        delete S;
        return bmi_fail(BMI_E_NOMEM, "server_create: arena of %zu bytes", cfg->hbm_budget);
    }
    S->arena.reset(S->cfg.hbm_budget);
    for (int i = 0; i < S->cfg.num_streams; ++i) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : S->streams) (void)hipStreamDestroy(x);
            (void)hipFree(S->base);
            delete S;
            return bmi_fail(BMI_E_HIP, "server_create: stream");
        }
        S->streams.push_back(s);
        S->stream_busy.push_back(false);
        // pinned staging for the plan tables, allocated here rather than on
        // the stream's first batch (a pinned allocation takes a few hundred
        // microseconds, which the first batches of a burst would wait for);
        // a batch with larger tables grows it
        void *h = nullptr;
        const bool fused = S->cfg.mode == BMI_SERVE_FUSED;
        if (fused && hipHostMalloc(&h, kStagingBytes, 0) != hipSuccess) h = nullptr;
        S->staging.push_back(h);
        S->staging_bytes.push_back(h ? kStagingBytes : 0);
    }
    S->t0 = std::chrono::steady_clock::now();
    S->thread = std::thread([S] { S->run(); });
    *out = S;
    return BMI_OK;
}

extern "C" int BMI_server_submit(BmiServer *S, const BmiRequest *req, unsigned long long *ticket)
{
    if (!S || !req || !ticket) return bmi_fail(BMI_E_INVALID, "server_submit: bad args");
    // ignored components are dropped at submit: the request is queued with
    // the remaining ones in the lower slots
    BmiRequest eff = *req;
    if (int rc = bmi_apply_ignore(req->tile, &eff.tile))
        return bmi_fail(rc, "server_submit: ignore mask");
    req = &eff;
    const BmiTileDesc &t = req->tile;
    if (t.out_format == IMAGE_FOTMAT_NONE || pixel_bytes(t.out_format) == 0)
        return bmi_fail(BMI_E_UNSUPPORTED, "server_submit: output format %d not served",
                        t.out_format);
    const double t_submit = now_us(S->t0);
    // validates the tile (0 = rejected) and bounds its plan's device bytes
    const size_t est = bmi_plan_estimate_bytes(&t, &S->cfg.tuning);
    if (!est) {
        BmiBatchPlan *P = nullptr;  // for the precise error code
        int rc = bmi_plan_build(&P, &t, 1, &S->cfg.tuning);
        if (P) BMI_batch_plan_destroy(P);
        return bmi_fail(rc ? rc : BMI_E_INVALID, "server_submit: request rejected");
    }
    Req *q = new Req();
    q->r = *req;
    output_size(t, &q->ow, &q->oh);
    q->dev_pitch = (t.width + 63) & ~63;
    const bool plain = S->cfg.mode == BMI_SERVE_PLAIN;
    q->plan_bytes = plain ? 0 : al(est);
    const size_t plane = al((size_t)q->dev_pitch * t.height * 4);
    if (req->host_io || plain) q->io_bytes += plane * t.ncomp;
    if (req->host_io) q->io_bytes += al((size_t)q->ow * q->oh * pixel_bytes(t.out_format));
    if (plain) q->io_bytes += al((size_t)t.width * t.height * 4 + 256);
    // the batch block is aligned: price the request as a batch of one
    if (q->plan_bytes + q->io_bytes + kAlign > S->cfg.hbm_budget) {
        delete q;
        return bmi_fail(BMI_E_NOMEM, "server_submit: footprint exceeds the HBM budget");
    }
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->stop) {
            delete q;
            return bmi_fail(BMI_E_INVALID, "server_submit: server is stopping");
        }
        q->ticket = S->next_ticket++;
        q->res.submit_us = t_submit;
        S->reqs[q->ticket] = q;
        S->pending.push_back(q);
        S->stats.submitted++;
    }
    S->cv_work.notify_one();
    *ticket = q->ticket;
    return BMI_OK;
}

extern "C" int BMI_server_wait(BmiServer *S, unsigned long long ticket, int timeout_ms,
                               BmiRequestResult *res)
{
    if (!S) return bmi_fail(BMI_E_INVALID, "server_wait: null server");
    std::unique_lock<std::mutex> g(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return bmi_fail(BMI_E_INVALID, "server_wait: unknown ticket");
    Req *q = it->second;
    auto ready = [&] { return q->done; };
    if (timeout_ms < 0)
        S->cv_done.wait(g, ready);
    else if (!S->cv_done.wait_for(g, std::chrono::milliseconds(timeout_ms), ready))
        return 1;
    if (res) *res = q->res;
    const int st = q->res.status;
    S->reqs.erase(it);
    delete q;
    return st;
}

extern "C" int BMI_server_stats(BmiServer *S, BmiServerStats *st)
{
    if (!S || !st) return bmi_fail(BMI_E_INVALID, "server_stats: bad args");
    std::lock_guard<std::mutex> g(S->mu);
    *st = S->stats;
    return BMI_OK;
}

extern "C" int BMI_server_destroy(BmiServer *S)
{
    if (!S) return BMI_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->thread.joinable()) S->thread.join();
    for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
    for (void *h : S->staging)
        if (h) (void)hipHostFree(h);
    if (S->base) (void)hipFree(S->base);
    for (auto &kv : S->reqs) delete kv.second;
    delete S;
    return BMI_OK;
}
