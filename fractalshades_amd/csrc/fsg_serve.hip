// fsg_serve.hip -- the serving loop: tile requests of mixed size, zoom depth
// and max_iter, admitted as device memory frees up and continuously batched
// onto the GPU.
//
// Design.  One scheduler thread per server owns L "lanes" (a HIP stream, a
// completion event and a small work buffer each) and one copy stream.  A
// request is split into chunks of `chunk_px` pixels; every chunk is one
// launch of the persistent tile kernel on a free lane, capped at 1/L of the
// GPU's resident waves so that up to L chunks -- from different requests,
// contexts, models and precisions -- run on the device at once.  When a lane
// frees up, it takes a chunk of the admitted request that has had the fewest
// chunks dispatched so far (then the oldest): a request that arrives while a
// large one is running gets the next free lane instead of waiting for the
// large one to drain.
//
// Admission is by device-memory footprint: a request needs its pixel offsets
// and its full result arrays on the device (16 + 16 n_Z + 9 bytes per pixel
// for the holomorphic models).  These are carved out of one arena of
// `hbm_budget` bytes allocated at start-up; pending requests are admitted in
// arrival order whenever the arena can hold the next one.  The reference
// data (orbit, derivative paths, BLA tree) stays resident in the contexts,
// which the caller keeps alive.
//
// Each request's results are the same whether it runs alone or in a mix:
// a pixel's computation depends only on its context and offset, never on the
// chunking or on what else runs.
//
// No reference counterpart: the reference renders tiles of one frame with a
// thread pool (mthreading.py:48-68) and has no request loop.
#include "fsgpu.h"
#include "fsg_internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace fsg {
int set_error(int code, const std::string &msg);
int ctx_holomorphic(const fsg_ctx *c);
int ctx_nz(const fsg_ctx *c);
size_t ctx_z_bytes(const fsg_ctx *c, int64_t n);
int ctx_launch_range(fsg_ctx *c, const double *pix, int64_t n, int64_t ld, double *Z, int32_t *U,
                     int8_t *sr, int32_t *si, unsigned long long *work, hipStream_t stream, int max_blocks,
                     int64_t grid_w);
size_t work_words();
int post_launch(const fsg_post_desc *pp, const double *Z, int64_t n, int64_t ld, const double *pix,
                const int32_t *si, const int8_t *sr, float *nu, float *dem, float *nrm, hipStream_t stream);
}  // namespace fsg

using namespace fsg;

namespace {

int64_t al256(int64_t x) { return (x + 255) & ~(int64_t)255; }

// first-fit allocator over [0, size) with 256-byte granules
class Arena {
public:
    void reset(int64_t size) { free_.clear(); if (size > 0) free_[0] = size; size_ = size; }
    int64_t size() const { return size_; }
    int64_t alloc(int64_t bytes) {
        bytes = al256(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const int64_t off = it->first, rest = it->second - bytes;
            free_.erase(it);
            if (rest > 0) free_[off + bytes] = rest;
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = al256(bytes);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t free_bytes() const {
        int64_t s = 0;
        for (auto &kv : free_) s += kv.second;
        return s;
    }
private:
    std::map<int64_t, int64_t> free_;
    int64_t size_ = 0;
};

struct Layout { int64_t pix, z, u, si, sr, nu, dem, nrm, total; };

Layout layout_of(const fsg_ctx *c, int64_t n, bool layers = false) {
    Layout l;
    l.pix = 0;
    l.z = l.pix + al256(16 * n);
    l.u = l.z + al256((int64_t)ctx_z_bytes(c, n));
    l.si = l.u + al256(4 * n);
    l.sr = l.si + al256(4 * n);
    l.nu = l.sr + al256(n);
    l.dem = l.nu + (layers ? al256(4 * n) : 0);
    l.nrm = l.dem + (layers ? al256(4 * n) : 0);
    l.total = l.nrm + (layers ? al256(8 * n) : 0);
    return l;
}

enum State { PENDING = 0, ACTIVE = 1, COPYING = 2, DONE = 3 };

struct Req {
    uint64_t id = 0;
    fsg_ctx *ctx = nullptr;
    const double *pix_h = nullptr;
    double *Z_h = nullptr; int32_t *U_h = nullptr; int8_t *sr_h = nullptr; int32_t *si_h = nullptr;
    bool layers = false;                  // return nu / dem instead of Z / U
    fsg_post_desc pp{};
    float *nu_h = nullptr, *dem_h = nullptr, *nrm_h = nullptr;
    int64_t n = 0;
    int64_t grid_w = 0;       // > 0: the pixels are rows of grid_w, each chunk taken in 8x8 blocks
    Layout lay{};
    int64_t off = -1;
    char *dev = nullptr;
    int32_t chunks = 0, dispatched = 0, finished = 0;
    hipEvent_t h2d = nullptr, d2h = nullptr;
    int state = PENDING;
    int status = 0;
    int waiters = 0;          // threads inside fsg_server_wait on this request
    bool released = false;    // its ticket has been returned by a wait
    std::string msg;   // error message of a failed request (the scheduler's thread-local one)
    double t_submit = 0, t_admit = 0, t_first = -1, t_done = 0;
};

struct Lane {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    unsigned long long *work = nullptr;
    Req *req = nullptr;
};

}  // namespace

struct fsg_server {
    fsg_server_opts o{};
    int device = 0;
    char *base = nullptr;
    Arena arena;
    hipStream_t copy = nullptr;
    std::vector<Lane> lanes;
    std::vector<hipEvent_t> ev_pool;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::deque<Req *> pending;                  // guarded by mu
    std::unordered_map<uint64_t, Req *> reqs;   // guarded by mu
    uint64_t next_id = 1;
    std::vector<Req *> active;                  // scheduler thread only
    std::thread th;
    std::chrono::steady_clock::time_point t0;

    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    hipEvent_t take_event() {
        if (!ev_pool.empty()) { hipEvent_t e = ev_pool.back(); ev_pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        return e;
    }
    void give_event(hipEvent_t e) { if (e) ev_pool.push_back(e); }

    void loop();
    void admit(Req *r);
    void finish(Req *r, int status);
};

// record a failure of r (first one wins) with its message
int fail_req(Req *r, int code, const char *msg) {
    if (r->status == 0) { r->status = code; r->msg = msg; set_error(code, msg); }
    return r->status;
}

void fsg_server::finish(Req *r, int status) {
    if (r->off >= 0) { arena.release(r->off, r->lay.total); r->off = -1; }
    give_event(r->h2d); give_event(r->d2h);
    r->h2d = r->d2h = nullptr;
    std::lock_guard<std::mutex> g(mu);
    r->status = status;
    r->t_done = now();
    r->state = DONE;
    cv_done.notify_all();
}

void fsg_server::admit(Req *r) {
    r->dev = base + r->off;
    r->t_admit = now();
    r->h2d = take_event();
    r->d2h = take_event();
    if (!r->h2d || !r->d2h) { finish(r, fail_req(r, FSG_E_HIP, "event creation failed")); return; }
    if (hipMemcpyAsync(r->dev + r->lay.pix, r->pix_h, (size_t)(16 * r->n), hipMemcpyHostToDevice, copy) != hipSuccess ||
        hipEventRecord(r->h2d, copy) != hipSuccess) {
        finish(r, fail_req(r, FSG_E_HIP, "input copy failed"));
        return;
    }
    r->chunks = (int32_t)((r->n + o.chunk_px - 1) / o.chunk_px);
    r->state = ACTIVE;
    active.push_back(r);
}

void fsg_server::loop() {
    if (hipSetDevice(device) != hipSuccess) return;
    for (;;) {
        bool progress = false;
        std::vector<Req *> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (stop && pending.empty() && active.empty()) break;
            while (!pending.empty()) {
                Req *r = pending.front();
                const int64_t off = arena.alloc(r->lay.total);
                if (off < 0) break;
                r->off = off;
                pending.pop_front();
                admitted.push_back(r);
            }
        }
        for (Req *r : admitted) { admit(r); progress = true; }

        // chunks that completed
        for (Lane &ln : lanes) {
            if (!ln.req) continue;
            const hipError_t q = hipEventQuery(ln.ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) fail_req(ln.req, FSG_E_HIP, "tile kernel failed");
            ln.req->finished++;
            ln.req = nullptr;
            progress = true;
        }
        // requests whose chunks are all done: results back to the host
        for (size_t i = 0; i < active.size();) {
            Req *r = active[i];
            if (r->state == ACTIVE && r->finished == r->chunks) {
                const int64_t n = r->n;
                bool ok = r->status == 0;
                if (ok && n > 0 && r->layers) {
                    float *nu = r->nu_h ? reinterpret_cast<float *>(r->dev + r->lay.nu) : nullptr;
                    float *dem = r->dem_h ? reinterpret_cast<float *>(r->dev + r->lay.dem) : nullptr;
                    float *nrm = r->nrm_h ? reinterpret_cast<float *>(r->dev + r->lay.nrm) : nullptr;
                    const int prc = post_launch(&r->pp, reinterpret_cast<const double *>(r->dev + r->lay.z), n, n,
                                                reinterpret_cast<const double *>(r->dev + r->lay.pix),
                                                reinterpret_cast<const int32_t *>(r->dev + r->lay.si),
                                                reinterpret_cast<const int8_t *>(r->dev + r->lay.sr), nu, dem, nrm,
                                                copy);
                    if (prc) { fail_req(r, prc, fsg_last_error()); ok = false; }
                    if (ok && nu) ok = hipMemcpyAsync(r->nu_h, nu, (size_t)(4 * n), hipMemcpyDeviceToHost, copy) == hipSuccess;
                    if (ok && dem) ok = hipMemcpyAsync(r->dem_h, dem, (size_t)(4 * n), hipMemcpyDeviceToHost, copy) == hipSuccess;
                    if (ok && nrm) ok = hipMemcpyAsync(r->nrm_h, nrm, (size_t)(8 * n), hipMemcpyDeviceToHost, copy) == hipSuccess;
                    ok = ok && hipMemcpyAsync(r->si_h, r->dev + r->lay.si, (size_t)(4 * n), hipMemcpyDeviceToHost, copy) == hipSuccess &&
                         hipMemcpyAsync(r->sr_h, r->dev + r->lay.sr, (size_t)n, hipMemcpyDeviceToHost, copy) == hipSuccess;
                    if (!ok) fail_req(r, FSG_E_HIP, "layer copy failed");
                } else if (ok && n > 0) {
                    ok = hipMemcpyAsync(r->Z_h, r->dev + r->lay.z, ctx_z_bytes(r->ctx, n), hipMemcpyDeviceToHost, copy) == hipSuccess &&
                         hipMemcpyAsync(r->U_h, r->dev + r->lay.u, (size_t)(4 * n), hipMemcpyDeviceToHost, copy) == hipSuccess &&
                         hipMemcpyAsync(r->si_h, r->dev + r->lay.si, (size_t)(4 * n), hipMemcpyDeviceToHost, copy) == hipSuccess &&
                         hipMemcpyAsync(r->sr_h, r->dev + r->lay.sr, (size_t)n, hipMemcpyDeviceToHost, copy) == hipSuccess;
                    if (!ok) fail_req(r, FSG_E_HIP, "result copy failed");
                }
                if (hipEventRecord(r->d2h, copy) != hipSuccess) fail_req(r, FSG_E_HIP, "event record failed");
                r->state = COPYING;
                progress = true;
            }
            if (r->state == COPYING) {
                const hipError_t q = hipEventQuery(r->d2h);
                if (q != hipErrorNotReady) {
                    if (q != hipSuccess) fail_req(r, FSG_E_HIP, "result copy failed");
                    active.erase(active.begin() + (long)i);
                    finish(r, r->status);
                    progress = true;
                    continue;
                }
            }
            ++i;
        }
        // free lanes take the next chunk: fewest chunks dispatched first
        for (Lane &ln : lanes) {
            if (ln.req) continue;
            Req *best = nullptr;
            for (Req *r : active) {
                if (r->state != ACTIVE || r->dispatched >= r->chunks) continue;
                if (!best || r->dispatched < best->dispatched ||
                    (r->dispatched == best->dispatched && r->id < best->id))
                    best = r;
            }
            if (!best) break;
            const int64_t start = (int64_t)best->dispatched * o.chunk_px;
            const int64_t cnt = std::min<int64_t>(o.chunk_px, best->n - start);
            best->dispatched++;
            if (best->t_first < 0) best->t_first = now();
            const int zw = ctx_holomorphic(best->ctx) ? 2 : 1;   // doubles per Z element
            char *d = best->dev;
            int rc = best->status;
            if (rc == 0 && hipStreamWaitEvent(ln.s, best->h2d, 0) != hipSuccess)
                rc = fail_req(best, FSG_E_HIP, "stream wait failed");
            if (rc == 0)
                rc = ctx_launch_range(best->ctx, reinterpret_cast<double *>(d + best->lay.pix) + 2 * start, cnt, best->n,
                                      reinterpret_cast<double *>(d + best->lay.z) + zw * start,
                                      reinterpret_cast<int32_t *>(d + best->lay.u) + start,
                                      reinterpret_cast<int8_t *>(d + best->lay.sr) + start,
                                      reinterpret_cast<int32_t *>(d + best->lay.si) + start,
                                      ln.work, ln.s, -(int)lanes.size(), best->grid_w);
            if (rc != 0) fail_req(best, rc, fsg_last_error());
            (void)hipEventRecord(ln.ev, ln.s);
            ln.req = best;
            progress = true;
        }
        if (!progress) {
            std::unique_lock<std::mutex> lk(mu);
            bool busy = !active.empty();
            if (busy)
                cv_work.wait_for(lk, std::chrono::microseconds(40));
            else
                cv_work.wait_for(lk, std::chrono::milliseconds(5), [&] { return stop || !pending.empty(); });
        }
    }
}

extern "C" {

int fsg_server_create(const fsg_server_opts *opts, fsg_server **out) {
    if (!out) return set_error(FSG_E_INVALID, "null output pointer");
    *out = nullptr;
    fsg_server_opts o{};
    if (opts) o = *opts;
    if (o.lanes == 0) o.lanes = 3;
    if (o.chunk_px == 0) o.chunk_px = 65536;
    if (o.hbm_budget == 0) o.hbm_budget = (int64_t)4 << 30;
    if (o.lanes < 1 || o.lanes > 16) return set_error(FSG_E_INVALID, "lanes must be in 1..16");
    if (o.chunk_px < 256) return set_error(FSG_E_INVALID, "chunk_px must be >= 256");
    if (o.hbm_budget < (1 << 20)) return set_error(FSG_E_INVALID, "hbm_budget must be >= 1 MiB");
    o.chunk_px = (o.chunk_px + 255) / 256 * 256;
    fsg_server *s = new (std::nothrow) fsg_server();
    if (!s) return set_error(FSG_E_NOMEM, "out of host memory");
    s->o = o;
    s->t0 = std::chrono::steady_clock::now();
    auto cleanup = [&](int rc) {
        for (Lane &ln : s->lanes) {
            if (ln.work) (void)hipFree(ln.work);
            if (ln.ev) (void)hipEventDestroy(ln.ev);
            if (ln.s) (void)hipStreamDestroy(ln.s);
        }
        if (s->copy) (void)hipStreamDestroy(s->copy);
        if (s->base) (void)hipFree(s->base);
        delete s;
        return rc;
    };
    if (hipGetDevice(&s->device) != hipSuccess) return cleanup(set_error(FSG_E_HIP, "no HIP device"));
    if (hipMalloc(reinterpret_cast<void **>(&s->base), (size_t)o.hbm_budget) != hipSuccess)
        return cleanup(set_error(FSG_E_NOMEM, "hipMalloc of the request arena failed"));
    s->arena.reset(o.hbm_budget);
    if (hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess)
        return cleanup(set_error(FSG_E_HIP, "stream creation failed"));
    s->lanes.resize((size_t)o.lanes);
    for (Lane &ln : s->lanes) {
        if (hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ln.ev, hipEventDisableTiming) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&ln.work), work_words() * sizeof(unsigned long long)) != hipSuccess)
            return cleanup(set_error(FSG_E_HIP, "lane creation failed"));
    }
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return 0;
}

int fsg_server_destroy(fsg_server *s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    for (Lane &ln : s->lanes) {
        (void)hipStreamSynchronize(ln.s);
        (void)hipFree(ln.work);
        (void)hipEventDestroy(ln.ev);
        (void)hipStreamDestroy(ln.s);
    }
    (void)hipStreamSynchronize(s->copy);
    (void)hipStreamDestroy(s->copy);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    (void)hipFree(s->base);
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
    return 0;
}

int fsg_server_footprint(const fsg_ctx *ctx, int64_t n, int64_t *bytes) {
    if (!ctx || !bytes || n < 0) return set_error(FSG_E_INVALID, "bad argument");
    *bytes = layout_of(ctx, n).total;
    return 0;
}

int fsg_server_clock(fsg_server *s, double *ms) {
    if (!s || !ms) return set_error(FSG_E_INVALID, "null argument");
    *ms = s->now();
    return 0;
}

}  // extern "C"

static int submit_req(fsg_server *s, Req *r, uint64_t *ticket);

extern "C" {

int fsg_server_submit(fsg_server *s, fsg_ctx *ctx, const double *c_pix, int64_t n, int64_t grid_w, double *Z,
                      int32_t *U, int8_t *stop_reason, int32_t *stop_iter, uint64_t *ticket) {
    if (!s || !ctx || !ticket) return set_error(FSG_E_INVALID, "null argument");
    if (n < 0 || (n > 0 && (!c_pix || !Z || !U || !stop_reason || !stop_iter)))
        return set_error(FSG_E_INVALID, "bad tile buffers");
    if (grid_w < 0) return set_error(FSG_E_INVALID, "grid_w must be >= 0");
    Req *r = new (std::nothrow) Req();
    if (!r) return set_error(FSG_E_NOMEM, "out of host memory");
    r->ctx = ctx; r->pix_h = c_pix; r->n = n; r->grid_w = grid_w;
    r->Z_h = Z; r->U_h = U; r->sr_h = stop_reason; r->si_h = stop_iter;
    r->lay = layout_of(ctx, n);
    return submit_req(s, r, ticket);
}

int fsg_server_submit_layers(fsg_server *s, fsg_ctx *ctx, const double *c_pix, int64_t n, int64_t grid_w,
                             const fsg_post_desc *pp, float *nu, float *dem, float *normal,
                             int8_t *stop_reason, int32_t *stop_iter, uint64_t *ticket) {
    if (!s || !ctx || !ticket || !pp) return set_error(FSG_E_INVALID, "null argument");
    if (n < 0 || (n > 0 && (!c_pix || !stop_reason || !stop_iter)))
        return set_error(FSG_E_INVALID, "bad tile buffers");
    if (grid_w < 0) return set_error(FSG_E_INVALID, "grid_w must be >= 0");
    if (!(pp->d > 1.) || !(pp->M > 0.) || ((dem || normal) && pp->dzndc_row < 1) ||
        pp->dzndc_row >= ctx_nz(ctx) || pp->holomorphic != ctx_holomorphic(ctx))
        return set_error(FSG_E_INVALID, "post-processing descriptor does not match the context");
    Req *r = new (std::nothrow) Req();
    if (!r) return set_error(FSG_E_NOMEM, "out of host memory");
    r->ctx = ctx; r->pix_h = c_pix; r->n = n; r->grid_w = grid_w;
    r->sr_h = stop_reason; r->si_h = stop_iter;
    r->layers = true; r->pp = *pp; r->nu_h = nu; r->dem_h = dem; r->nrm_h = normal;
    r->lay = layout_of(ctx, n, true);
    return submit_req(s, r, ticket);
}

}  // extern "C"

static int submit_req(fsg_server *s, Req *r, uint64_t *ticket) {
    if (r->lay.total > s->arena.size()) {
        delete r;
        return set_error(FSG_E_NOMEM, "request footprint exceeds the server's device-memory budget");
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) { delete r; return set_error(FSG_E_INVALID, "server is stopping"); }
        r->id = s->next_id++;
        r->t_submit = s->now();
        s->reqs[r->id] = r;
        s->pending.push_back(r);
        *ticket = r->id;
    }
    s->cv_work.notify_all();
    return 0;
}

extern "C" {

int fsg_server_wait(fsg_server *s, uint64_t ticket, int32_t timeout_ms, fsg_request_info *info) {
    if (!s) return set_error(FSG_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_error(FSG_E_INVALID, "unknown ticket");
    Req *r = it->second;
    // several threads may wait on one ticket: the request lives until the
    // last of them leaves; the first to see it done releases the ticket
    r->waiters++;
    auto done = [&] { return r->state == DONE; };
    bool ok = true;
    if (timeout_ms < 0) s->cv_done.wait(lk, done);
    else ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
    r->waiters--;
    if (!ok) return FSG_PENDING;
    if (!r->released) {
        r->released = true;
        s->reqs.erase(ticket);
    }
    if (info) {
        info->status = r->status;
        info->chunks = r->chunks;
        info->footprint = r->lay.total;
        info->t_submit_ms = r->t_submit;
        info->t_admit_ms = r->t_admit;
        info->t_first_ms = r->t_first;
        info->t_done_ms = r->t_done;
    }
    const int st = r->status;
    if (st != 0) set_error(st, r->msg);
    if (r->waiters == 0) delete r;
    return st;
}

void *fsg_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        set_error(FSG_E_NOMEM, "hipHostMalloc failed");
        return nullptr;
    }
    return p;
}

int fsg_host_free(void *p) {
    if (p && hipHostFree(p) != hipSuccess) return set_error(FSG_E_HIP, "hipHostFree failed");
    return 0;
}

}  // extern "C"
