// orb_serve.cpp -- continuous-batching request loop for many camera streams.
//
// Each request is one frame of one stream: ORB extraction (the hot path of
// Frame's constructors) plus a brute-force Hamming match of the new
// descriptors against that stream's previous-frame descriptors, which stay
// resident in HBM.  A server thread runs the loop:
//   * completions: batches whose HIP event has fired are finalised, their
//     extractor slots released and their latencies stamped;
//   * admission: pending requests, oldest first, take a free extractor slot
//     of their size (allocating one while the HBM budget allows); a stream
//     contributes at most one frame per batch, so its frames stay ordered;
//   * launch: the admitted requests form one batch on a free HIP stream, so
//     several batches overlap on the device: one host-to-device copy of the
//     batch's tables, one launch that pulls every frame from its pinned
//     staging buffer into its slot's level 0, one launch per pipeline stage
//     over all of them, one multi-pair match launch (which also updates the
//     resident descriptors), and one launch that writes every request's
//     results straight into its pinned result buffer.  The HIP calls per
//     batch do not grow with its number of frames.
// Requests arriving while batches run are admitted as soon as a slot and a
// HIP stream free up: continuous batching rather than a fixed batch window.
// Frames of one stream may be in flight in several lanes at once (one per
// batch): their extraction overlaps, and only the match of a frame waits
// (hipStreamWaitEvent) for the batch that holds the stream's previous frame.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <random>
#include <thread>

#include "orb_internal.h"

namespace {

double clock_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Camera {
    int w, h, nfeatures, cap;
    uint8_t* d_prev = nullptr;   // previous frame's descriptors (cap x 32)
    int* d_prev_n = nullptr;
    int inflight = 0;            // frames of this stream on the device
    int last_lane = -1;          // lane holding its newest in-flight frame
    int affinity = -1;           // last slot used (keeps its pyramid resident)
};

struct Slot {
    orb_extractor* ex = nullptr;
    int w, h, nfeatures;
    int* d_midx = nullptr;
    int* d_mdist = nullptr;
    bool free = true;
    size_t bytes = 0;
};

struct Request {
    uint64_t ticket;
    int stream;
    double t_arrival, t_admit = 0, t_done = 0;
    uint8_t* h_img = nullptr;    // pinned staging copy of the frame, rows at the level-0 pitch
    char* h_res = nullptr;       // pinned results: counters | kps | desc | idx | dist
    int slot = -1, cap = 0, status = ORB_OK, n_kps = 0, n_matches = 0;
    int dep_lane = -1;           // lane whose batch holds this stream's previous frame
};

// Per-batch tables, in one pinned buffer and its device twin: jobs | run
// table | match args | input copies | output copies.
constexpr size_t kStageJobs = 0;
constexpr size_t kStageRt = kStageJobs + ORB_MAX_JOBS * sizeof(ImageJob);
constexpr size_t kStageMargs = kStageRt + sizeof(RunTable);
constexpr size_t kStageIn = (kStageMargs + ORB_MAX_JOBS * sizeof(MatchArgs) + 15) / 16 * 16;
constexpr int kOutOps = 5;       // counters, keypoints, descriptors, match index, match distance
constexpr size_t kStageOut = kStageIn + ORB_MAX_JOBS * sizeof(CopyOp);
constexpr size_t kStageBytes = kStageOut + ORB_MAX_JOBS * kOutOps * sizeof(CopyOp);

struct Lane {                    // one HIP stream with its staging
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    char* d_stage = nullptr;
    char* h_stage = nullptr;
    std::vector<Request*> reqs;
    bool busy = false;
    bool failed = false;         // the batch's launch failed: its results are not valid
};

size_t result_bytes(int cap) { return 256 + (size_t)cap * (sizeof(OrbKeyPoint) + 32 + 8); }

// Row pitch of a frame's pinned staging copy: that of the extractor's level 0.
size_t img_pitch(int w) { return ((size_t)w + 63) / 64 * 64; }

}  // namespace

struct orb_server {
    CtxRef ctx;
    float scaleFactor = 1.2f;
    int nlevels = 8, iniTh = 20, minTh = 7, th_low = 50;
    float ratio = 0.6f;
    size_t budget = 0, used = 0;
    int max_batch = 64;
    std::vector<Camera> cams;
    std::vector<Slot> slots;
    std::vector<Lane> lanes;
    size_t max_img = 0;
    int max_cap = 0;
    std::vector<uint8_t*> img_pool;
    std::vector<char*> res_pool;

    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request*> pending;
    std::vector<Request*> completed;
    std::map<uint64_t, Request*> done;
    uint64_t next_ticket = 1;
    std::atomic<bool> stop{false};
    std::thread th;
    int64_t n_batches = 0, n_requests = 0;
    int inflight = 0;
    std::atomic<int> error{ORB_OK};

    // Free slot for a frame of camera c: -1 = wait for one, -2 = the memory
    // budget can never hold a slot of this size.
    int slot_for(Camera& c)
    {
        if (c.affinity >= 0 && slots[c.affinity].free) return c.affinity;
        for (size_t i = 0; i < slots.size(); i++) {
            const Slot& s = slots[i];
            if (s.free && s.w == c.w && s.h == c.h && s.nfeatures == c.nfeatures) return (int)i;
        }
        // allocate a new slot if the memory budget allows it
        orb_extractor* e = nullptr;
        if (orb_extractor_create(ctx, c.nfeatures, scaleFactor, nlevels, iniTh, minTh, &e)) return -1;
        Plan p;
        if (e->make_plan(c.w, c.h, p)) {
            delete e;
            return -1;
        }
        const size_t need = p.img_bytes * 2 + p.ncells + p.nkeys * 20 + p.nret * 8 + (size_t)p.out_cap * 96 + 65536;
        bool any = false;
        for (const Slot& sl : slots) any |= sl.w == c.w && sl.h == c.h && sl.nfeatures == c.nfeatures;
        if (used + need > budget) {
            delete e;
            return any ? -1 : -2;
        }
        if (e->ensure(c.w, c.h)) {
            delete e;
            return -1;
        }
        Slot s;
        s.ex = e;
        s.w = c.w;
        s.h = c.h;
        s.nfeatures = c.nfeatures;
        s.bytes = e->mem_bytes + (size_t)e->plan.out_cap * 8;
        if (hipMalloc(&s.d_midx, (size_t)e->plan.out_cap * 8) != hipSuccess) {
            delete e;
            return -1;
        }
        s.d_mdist = s.d_midx + e->plan.out_cap;
        used += s.bytes;
        slots.push_back(s);
        return (int)slots.size() - 1;
    }

    void launch(Lane& L, std::vector<Request*>& batch)
    {
        const int n = (int)batch.size();
        ImageJob* jobs = (ImageJob*)(L.h_stage + kStageJobs);
        RunTable* rt = (RunTable*)(L.h_stage + kStageRt);
        MatchArgs* ma = (MatchArgs*)(L.h_stage + kStageMargs);
        CopyOp* in_ops = (CopyOp*)(L.h_stage + kStageIn);
        CopyOp* out_ops = (CopyOp*)(L.h_stage + kStageOut);
        std::vector<const orb_extractor*> exs(n);
        int waves = 0;
        for (int i = 0; i < n; i++) {
            Request* r = batch[i];
            Slot& s = slots[r->slot];
            Camera& c = cams[r->stream];
            const ImageJob& J = s.ex->hjob;
            exs[i] = s.ex;
            jobs[i] = J;
            // the frame, read by the device from its pinned staging copy
            in_ops[i] = CopyOp{(const unsigned*)r->h_img, (unsigned*)const_cast<uint8_t*>(J.src), nullptr,
                               (int)(img_pitch(c.w) * c.h / 4), 0};
            MatchArgs& m = ma[i];
            std::memset(&m, 0, sizeof(m));
            m.q = J.desc;
            m.nq = J.counters + ORB_CNT_NOUT;
            m.t = c.d_prev;
            m.nt = c.d_prev_n;
            m.th_low = th_low;
            m.ratio = ratio;
            m.idx = s.d_midx;
            m.dist = s.d_mdist;
            m.cap_q = s.ex->plan.out_cap;
            m.wave_base = waves;
            m.save_desc = c.d_prev;
            m.save_n = c.d_prev_n;
            m.save_cap = c.cap;
            waves += m.cap_q;
            // the results, written by the device into the pinned result
            // buffer: counters | keypoints | descriptors | index | distance
            const int cap = s.ex->plan.out_cap;
            const int* cnt = J.counters + ORB_CNT_NOUT;
            char* p = r->h_res;
            CopyOp* o = out_ops + (size_t)kOutOps * i;
            o[0] = CopyOp{(const unsigned*)J.counters, (unsigned*)p, nullptr, ORB_CNT_SLOTS, 0};
            p += 256;
            o[1] = CopyOp{(const unsigned*)J.kps, (unsigned*)p, cnt, cap * (int)(sizeof(OrbKeyPoint) / 4),
                          (int)(sizeof(OrbKeyPoint) / 4)};
            p += (size_t)cap * sizeof(OrbKeyPoint);
            o[2] = CopyOp{(const unsigned*)J.desc, (unsigned*)p, cnt, cap * 8, 8};
            p += (size_t)cap * 32;
            o[3] = CopyOp{(const unsigned*)s.d_midx, (unsigned*)p, cnt, cap, 1};
            o[4] = CopyOp{(const unsigned*)s.d_mdist, (unsigned*)(p + (size_t)cap * 4), cnt, cap, 1};
            r->cap = cap;
        }
        (void)orb_extractor::build_table(exs.data(), n, *rt);   // the server's extractors share tile heights
        hipError_t e = hipMemcpyAsync(L.d_stage, L.h_stage, kStageBytes, hipMemcpyHostToDevice, L.st);
        if (e == hipSuccess) e = orb_launch_copy_ops((const CopyOp*)(L.d_stage + kStageIn), n, 32, L.st);
        if (e == hipSuccess)
            e = orb_launch_extract((const ImageJob*)(L.d_stage + kStageJobs), (const RunTable*)(L.d_stage + kStageRt),
                                   rt, L.st, nullptr);
        // Extraction overlaps the previous frames of the same streams; the
        // match needs their descriptors, so it waits for those batches.
        std::vector<char> waited(lanes.size(), 0);
        for (Request* r : batch)
            if (r->dep_lane >= 0 && !waited[r->dep_lane]) {
                waited[r->dep_lane] = 1;
                if (e == hipSuccess) e = hipStreamWaitEvent(L.st, lanes[r->dep_lane].ev, 0);
            }
        if (e == hipSuccess) e = orb_launch_match_pairs((const MatchArgs*)(L.d_stage + kStageMargs), n, waves, L.st);
        if (e == hipSuccess)
            e = orb_launch_copy_ops((const CopyOp*)(L.d_stage + kStageOut), kOutOps * n, 4, L.st);
        if (e == hipSuccess) e = hipEventRecord(L.ev, L.st);
        L.failed = e != hipSuccess;
        if (L.failed) {
            error = ORB_E_DEVICE;
            (void)hipStreamSynchronize(L.st);   // whatever was enqueued has finished before finish() reuses it
        }
        L.reqs = batch;
    }

    void finish(Lane& L)
    {
        const double t = clock_s();
        std::lock_guard<std::mutex> g(mu);
        for (Request* r : L.reqs) {
            const int* ctr = (const int*)r->h_res;
            r->t_done = t;
            r->status = L.failed ? ORB_E_DEVICE : ctr[ORB_CNT_ERR] ? ORB_E_OVERFLOW : ORB_OK;
            r->n_kps = r->status ? 0 : ctr[ORB_CNT_NOUT];
            const int* idx = (const int*)(r->h_res + 256 + (size_t)r->cap * (sizeof(OrbKeyPoint) + 32));
            int m = 0;
            for (int i = 0; i < r->n_kps; i++) m += idx[i] >= 0;
            r->n_matches = m;
            slots[r->slot].free = true;
            Camera& c = cams[r->stream];
            c.inflight--;
            if (c.inflight == 0) c.last_lane = -1;
            c.affinity = r->slot;
            img_pool.push_back(r->h_img);
            r->h_img = nullptr;
            completed.push_back(r);
        }
        L.reqs.clear();
        L.busy = false;
        inflight--;
        cv.notify_all();
    }

    void loop()
    {
        (void)hipSetDevice(ctx->device);
        std::vector<Request*> batch;
        while (!stop.load()) {
            for (Lane& L : lanes)
                if (L.busy && hipEventQuery(L.ev) == hipSuccess) finish(L);
            Lane* free_lane = nullptr;
            for (Lane& L : lanes)
                if (!L.busy) {
                    free_lane = &L;
                    break;
                }
            batch.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                if (free_lane && !pending.empty()) {
                    std::vector<char> blocked(cams.size(), 0);
                    const double t = clock_s();
                    for (auto it = pending.begin(); it != pending.end() && (int)batch.size() < max_batch;) {
                        Request* r = *it;
                        Camera& c = cams[r->stream];
                        // one frame per stream per batch; at most one per lane in flight
                        if (blocked[r->stream] || c.inflight >= (int)lanes.size()) {
                            blocked[r->stream] = 1;
                            ++it;
                            continue;
                        }
                        const int s = slot_for(c);
                        if (s == -2) {
                            // refused: the budget cannot hold a slot of this size
                            r->status = ORB_E_ADMISSION;
                            r->t_admit = r->t_done = t;
                            img_pool.push_back(r->h_img);
                            r->h_img = nullptr;
                            completed.push_back(r);
                            it = pending.erase(it);
                            continue;
                        }
                        if (s < 0) {
                            blocked[r->stream] = 1;
                            ++it;
                            continue;
                        }
                        slots[s].free = false;
                        r->slot = s;
                        r->t_admit = t;
                        r->dep_lane = c.inflight > 0 ? c.last_lane : -1;
                        c.inflight++;
                        c.last_lane = (int)(free_lane - lanes.data());
                        blocked[r->stream] = 1;
                        batch.push_back(r);
                        it = pending.erase(it);
                    }
                }
                if (batch.empty()) {
                    if (inflight == 0) cv.wait_for(lk, std::chrono::milliseconds(1));
                } else {
                    // The batch counts as in flight from the moment it leaves
                    // `pending`, so drain() never sees it in neither place.
                    free_lane->busy = true;
                    inflight++;
                    n_batches++;
                    n_requests += (int64_t)batch.size();
                }
            }
            if (!batch.empty()) {
                launch(*free_lane, batch);
            } else if (inflight > 0) {
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    }

    ~orb_server()
    {
        stop.store(true);
        cv.notify_all();
        if (th.joinable()) th.join();
        (void)hipSetDevice(ctx->device);
        for (Lane& L : lanes) {
            if (L.st) (void)hipStreamSynchronize(L.st);
            if (L.ev) (void)hipEventDestroy(L.ev);
            if (L.st) (void)hipStreamDestroy(L.st);
            if (L.d_stage) (void)hipFree(L.d_stage);
            if (L.h_stage) (void)hipHostFree(L.h_stage);
            for (Request* r : L.reqs) pending.push_back(r);
        }
        for (Slot& s : slots) {
            delete s.ex;
            if (s.d_midx) (void)hipFree(s.d_midx);
        }
        for (Camera& c : cams)
            if (c.d_prev) (void)hipFree(c.d_prev);
        auto drop = [&](Request* r) {
            if (r->h_img) img_pool.push_back(r->h_img);
            if (r->h_res) res_pool.push_back(r->h_res);
            delete r;
        };
        for (Request* r : pending) drop(r);
        for (Request* r : completed) drop(r);
        for (auto& kv : done) drop(kv.second);
        for (uint8_t* p : img_pool) (void)hipHostFree(p);
        for (char* p : res_pool) (void)hipHostFree(p);
    }
};

extern "C" {

double orb_server_clock(void) { return clock_s(); }

int orb_server_create(orb_context* ctx, int nstreams, const int* widths, const int* heights, const int* nfeatures,
                      float scaleFactor, int nlevels, int iniThFAST, int minThFAST, int th_low, float ratio,
                      size_t hbm_budget, int max_inflight, int max_batch, orb_server** out)
{
    if (!ctx || !out || nstreams <= 0 || !widths || !heights || !nfeatures || max_inflight <= 0 ||
        max_inflight > 16 || max_batch <= 0 || max_batch > ORB_MAX_JOBS || nlevels < 1 || nlevels > ORB_MAX_LEVELS)
        return ORB_E_ARG;
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    orb_server* s = new orb_server();
    s->ctx = ctx;
    s->scaleFactor = scaleFactor;
    s->nlevels = nlevels;
    s->iniTh = iniThFAST;
    s->minTh = minThFAST;
    s->th_low = th_low;
    s->ratio = ratio;
    s->budget = hbm_budget;
    s->max_batch = max_batch;
    for (int i = 0; i < nstreams; i++) {
        Camera c;
        c.w = widths[i];
        c.h = heights[i];
        c.nfeatures = nfeatures[i];
        orb_extractor* probe = nullptr;
        if (orb_extractor_create(ctx, c.nfeatures, scaleFactor, nlevels, iniThFAST, minThFAST, &probe)) {
            delete s;
            return ORB_E_ARG;
        }
        const int cap = orb_extractor_capacity(probe, c.w, c.h);
        delete probe;
        if (cap <= 0) {
            delete s;
            return ORB_E_ARG;
        }
        c.cap = cap;
        if (hipMalloc(&c.d_prev, (size_t)cap * 32 + 256) != hipSuccess) {
            delete s;
            return ORB_E_OOM;
        }
        c.d_prev_n = (int*)(c.d_prev + (size_t)cap * 32);
        if (hipMemset(c.d_prev_n, 0, sizeof(int)) != hipSuccess) {
            delete s;
            return ORB_E_DEVICE;
        }
        s->used += (size_t)cap * 32 + 256;
        s->max_img = std::max(s->max_img, img_pitch(c.w) * c.h);
        s->max_cap = std::max(s->max_cap, cap);
        s->cams.push_back(c);
    }
    if (s->used > s->budget) {
        delete s;
        return ORB_E_ADMISSION;
    }
    s->lanes.resize(max_inflight);
    for (Lane& L : s->lanes) {
        if (hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&L.ev, hipEventDisableTiming | hipEventReleaseToSystem) != hipSuccess ||
            hipMalloc(&L.d_stage, kStageBytes) != hipSuccess || hipHostMalloc(&L.h_stage, kStageBytes) != hipSuccess) {
            delete s;
            return ORB_E_OOM;
        }
    }
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return ORB_OK;
}

void orb_server_destroy(orb_server* s) { delete s; }

int orb_server_submit(orb_server* s, int stream, const uint8_t* image, size_t stride, double t_arrival,
                      uint64_t* ticket)
{
    if (!s || stream < 0 || stream >= (int)s->cams.size() || !image) return ORB_E_ARG;
    const Camera& c = s->cams[stream];
    if (stride < (size_t)c.w) return ORB_E_ARG;
    uint8_t* img = nullptr;
    char* res = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->error.load()) return s->error.load();
        if (s->pending.size() >= 4096) return ORB_E_ADMISSION;
        if (!s->img_pool.empty()) {
            img = s->img_pool.back();
            s->img_pool.pop_back();
        }
        if (!s->res_pool.empty()) {
            res = s->res_pool.back();
            s->res_pool.pop_back();
        }
    }
    // Coherent (fine-grained) pinned memory: the device reads the frame and
    // writes the results over PCIe with no copy engine in between, so neither
    // side may keep them in a cache.
    if (!img && hipHostMalloc(&img, s->max_img, hipHostMallocCoherent) != hipSuccess) return ORB_E_OOM;
    if (!res && hipHostMalloc(&res, result_bytes(s->max_cap), hipHostMallocCoherent) != hipSuccess) return ORB_E_OOM;
    const size_t pitch = img_pitch(c.w);
    for (int y = 0; y < c.h; y++) std::memcpy(img + (size_t)y * pitch, image + (size_t)y * stride, c.w);
    Request* r = new Request();
    r->stream = stream;
    r->t_arrival = t_arrival;
    r->h_img = img;
    r->h_res = res;
    {
        std::lock_guard<std::mutex> g(s->mu);
        r->ticket = s->next_ticket++;
        s->pending.push_back(r);
    }
    s->cv.notify_all();
    if (ticket) *ticket = r->ticket;
    return ORB_OK;
}

int orb_server_poll(orb_server* s, orb_serve_info* out, int max, int* n)
{
    if (!s || !n || (max > 0 && !out)) return ORB_E_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    const int k_max = std::min(max, (int)s->completed.size());
    int k = 0;
    while (k < k_max) {
        Request* r = s->completed[k];
        orb_serve_info& o = out[k++];
        o.ticket = r->ticket;
        o.stream = r->stream;
        o.status = r->status;
        o.n_keypoints = r->n_kps;
        o.n_matches = r->n_matches;
        o.t_arrival = r->t_arrival;
        o.t_admit = r->t_admit;
        o.t_done = r->t_done;
        s->done[r->ticket] = r;
    }
    s->completed.erase(s->completed.begin(), s->completed.begin() + k);
    *n = k;
    return s->error.load();
}

int orb_server_fetch(orb_server* s, uint64_t ticket, orb_keypoint* kps, int cap, uint8_t* desc, int32_t* match_idx,
                     int32_t* match_dist, int* n)
{
    if (!s || !n) return ORB_E_ARG;
    Request* r = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->done.find(ticket);
        if (it == s->done.end()) return ORB_E_ARG;
        r = it->second;
        s->done.erase(it);
    }
    *n = r->n_kps;
    int rc = r->status;
    if (!rc && r->n_kps > cap && (kps || desc || match_idx || match_dist)) rc = ORB_E_CAPACITY;
    if (!rc) {
        const char* p = r->h_res + 256;
        const size_t k = (size_t)r->n_kps;
        if (kps) std::memcpy(kps, p, k * sizeof(OrbKeyPoint));
        p += (size_t)r->cap * sizeof(OrbKeyPoint);
        if (desc) std::memcpy(desc, p, k * 32);
        p += (size_t)r->cap * 32;
        if (match_idx) std::memcpy(match_idx, p, k * 4);
        if (match_dist) std::memcpy(match_dist, p + (size_t)r->cap * 4, k * 4);
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->res_pool.push_back(r->h_res);
    }
    delete r;
    return rc;
}

// Release a completed request without reading its results.
int orb_server_release(orb_server* s, uint64_t ticket)
{
    int n = 0;
    const int rc = orb_server_fetch(s, ticket, nullptr, 0, nullptr, nullptr, nullptr, &n);
    return rc == ORB_E_OVERFLOW ? ORB_OK : rc;
}

int orb_server_drain(orb_server* s, double timeout_s)
{
    if (!s) return ORB_E_ARG;
    const double t_end = clock_s() + timeout_s;
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->pending.empty() || s->inflight > 0) {
        if (clock_s() > t_end) return ORB_E_DEVICE;
        s->cv.wait_for(lk, std::chrono::milliseconds(1));
    }
    return s->error.load();
}

// Open-loop replay from native threads: one submitter per stream, the
// calling thread polling and releasing.  Arrival times are exponential
// inter-arrivals at fps per stream (one generator per stream, seeded with
// seed + stream), starting 50 ms after the call; frame k of a stream is
// frames[s][j] with j running 0 .. n-1 .. 1 0 1 .., so each match pairs
// sequence neighbours.  Latency = completion - scheduled arrival.
int orb_server_replay(orb_server* s, int nstreams, const uint8_t* const* frames, const int* nframes, double fps,
                      double duration, uint64_t seed, orb_replay_stats* out)
{
    if (!s || !out || nstreams <= 0 || nstreams > (int)s->cams.size() || !frames || !nframes || !(fps > 0) ||
        !(duration > 0))
        return ORB_E_ARG;
    for (int i = 0; i < nstreams; i++)
        if (!frames[i] || nframes[i] <= 0) return ORB_E_ARG;
    std::memset(out, 0, sizeof(*out));
    std::atomic<int64_t> submitted{0}, refused{0};
    std::atomic<int> failed{ORB_OK}, running{nstreams};
    const double t0 = clock_s() + 0.05;
    std::vector<std::thread> subs;
    for (int si = 0; si < nstreams; si++)
        subs.emplace_back([&, si]() {
            const Camera& c = s->cams[si];
            const size_t bytes = (size_t)c.w * c.h;
            const int n = nframes[si];
            std::mt19937_64 rng(seed + (uint64_t)si);
            std::exponential_distribution<double> gap(fps);
            double t = gap(rng);
            for (long k = 0; t < duration && failed.load() == ORB_OK; k++, t += gap(rng)) {
                const double at = t0 + t;
                for (double now = clock_s(); now < at; now = clock_s()) {
                    if (at - now > 3e-4)
                        std::this_thread::sleep_for(std::chrono::duration<double>(at - now - 2e-4));
                    else
                        std::this_thread::yield();
                }
                const long j = n <= 1 ? 0 : k % (2 * n - 2);
                const long f = j < n ? j : 2 * n - 2 - j;
                uint64_t ticket = 0;
                const int rc = orb_server_submit(s, si, frames[si] + (size_t)f * bytes, (size_t)c.w, at, &ticket);
                if (rc == ORB_E_ADMISSION) {
                    refused++;
                } else if (rc) {
                    failed = rc;
                } else {
                    submitted++;
                }
            }
            running--;
        });
    std::vector<double> lat;
    std::vector<orb_serve_info> info(1024);
    int64_t done = 0, errors = 0;
    double first = 1e300, last = 0, t_progress = clock_s();
    int rc = ORB_OK;
    while (true) {
        int k = 0;
        rc = orb_server_poll(s, info.data(), (int)info.size(), &k);
        if (rc) break;
        for (int i = 0; i < k; i++) {
            const orb_serve_info& o = info[i];
            lat.push_back((o.t_done - o.t_arrival) * 1e3);
            first = std::min(first, o.t_arrival);
            last = std::max(last, o.t_done);
            errors += o.status != ORB_OK;
            (void)orb_server_release(s, o.ticket);
        }
        done += k;
        if (k) t_progress = clock_s();
        if (running.load() == 0 && done >= submitted.load()) break;
        if (clock_s() - t_progress > 30.0) {
            rc = ORB_E_DEVICE;
            break;
        }
        if (!k) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    if (rc) failed = rc;
    for (std::thread& th : subs) th.join();
    if (failed.load()) return failed.load();
    out->requests = done;
    out->refused = refused.load();
    out->errors = errors;
    if (!lat.empty()) {
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) { return lat[std::min(lat.size() - 1, (size_t)(q * (double)(lat.size() - 1) + 0.5))]; };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        out->max_ms = lat.back();
        out->frames_per_s = last > first ? (double)done / (last - first) : 0.0;
    }
    return ORB_OK;
}

int orb_server_stats(orb_server* s, int64_t* batches, int64_t* requests, size_t* hbm_used, int* slots)
{
    if (!s) return ORB_E_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    if (batches) *batches = s->n_batches;
    if (requests) *requests = s->n_requests;
    if (hbm_used) *hbm_used = s->used;
    if (slots) *slots = (int)s->slots.size();
    return ORB_OK;
}

}  // extern "C"
