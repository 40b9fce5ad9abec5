// orb_internal.h -- host-side state shared by the runtime translation units
// (extractor handles, the context, launch entry points of orb_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "orb_gpu.h"
#include "orb_types.h"

extern "C" hipError_t orb_upload_constants(const int* umax16, const int* gauss7);
extern "C" hipError_t orb_launch_extract(const ImageJob* d_jobs, const RunTable* d_rt, const RunTable* h_rt,
                                         hipStream_t st, hipEvent_t* ev);
extern "C" hipError_t orb_launch_stereo(const StereoArgs* d_args, int npairs, int max_left, hipStream_t st,
                                        hipEvent_t* ev);
extern "C" hipError_t orb_launch_match_bf(const uint8_t* q, const int* nq, const uint8_t* t, const int* nt,
                                          int th_low, float ratio, int* idx, int* dist, int cap_q,
                                          hipStream_t st);

extern "C" hipError_t orb_launch_match_pairs(const MatchArgs* d_args, int npairs, int nwaves, hipStream_t st);
extern "C" hipError_t orb_launch_gray(const uint8_t* src, size_t sstride, int scn, int bidx, int w, int h,
                                      uint8_t* dst, size_t dstride, hipStream_t st);
extern "C" hipError_t orb_launch_distinctive(const uint8_t* desc, const int* offsets, int n_points, int* best,
                                             uint8_t* out_desc, hipStream_t st);
extern "C" hipError_t orb_launch_hbm_copy(const void* src, void* dst, size_t bytes, hipStream_t st);
extern "C" hipError_t orb_launch_copy_ops(const CopyOp* d_ops, int nops, int blocks_per_op, hipStream_t st);
extern "C" hipError_t orb_launch_undistort(const OrbKeyPoint* in, int n, const int* count, const double* K4,
                                           const double* k5, OrbKeyPoint* out, hipStream_t st);
extern "C" hipError_t orb_launch_rgbd_depth(const OrbKeyPoint* kps, const OrbKeyPoint* kun, int n, const int* count,
                                            const float* depth, int w, int h, size_t pitch, float bf, float* uR,
                                            float* dep, hipStream_t st);
extern "C" int orb_knn_chunks(long long nm);
extern "C" int orb_knn_scratch_setup(int device);
extern "C" uint2* orb_knn_scratch(int device, size_t* entries);
extern "C" hipError_t orb_launch_knn2(const uint8_t* q, int nq, const uint8_t* m, long long nm, int method,
                                      float ratio, uint2* part, int* idx2, int* dist2, uint8_t* accept,
                                      hipStream_t st, hipEvent_t* ev);

namespace orb_detail {

inline int cv_round(float v) { return (int)std::lrintf(v); }

#define HIP_TRY(x)                               \
    do {                                         \
        if ((x) != hipSuccess) return ORB_E_DEVICE; \
    } while (0)

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Integer taps of getGaussianKernel(7, 2) scaled by 256 (OpenCV's 8U
// separable path): float kernel normalised in double, rounded half-even.
inline void gaussian_taps(int k[7])
{
    float cf[7];
    double sum = 0;
    for (int i = 0; i < 7; i++) {
        const double x = i - 3.0;
        cf[i] = (float)std::exp((-0.5 / (2.0 * 2.0)) * x * x);
        sum += cf[i];
    }
    sum = 1. / sum;
    for (int i = 0; i < 7; i++) cf[i] = (float)(cf[i] * sum);
    for (int i = 0; i < 7; i++) k[i] = cv_round(cf[i] * 256.f);
}

// umax: half-widths of the radius-15 orientation disc (ORBextractor.cc:542-570).
inline void disc_umax(int umax[16])
{
    const int half = 15;
    const int vmax = (int)std::floor(half * std::sqrt(2.f) / 2 + 1);
    const int vmin = (int)std::ceil(half * std::sqrt(2.f) / 2);
    const double hp2 = half * half;
    for (int v = 0; v <= vmax; ++v) umax[v] = (int)std::lrint(std::sqrt(hp2 - v * v));
    for (int v = half, v0 = 0; v >= vmin; --v) {
        while (umax[v0] == umax[v0 + 1]) ++v0;
        umax[v] = v0;
        ++v0;
    }
}

// Border index of REFLECT_101 (the level kernel's halo rule).
// First level column that OpenCV's SSE2 VResizeLinearVec_32s8u leaves to
// the scalar pass: it covers 16 columns at a time while x <= w - 16, then 4
// at a time while x < w - 4 (imgwarp.cpp, OpenCV 2.4/3.x).
inline int sse2_resize_end(int w)
{
    int x = 0;
    while (x <= w - 16) x += 16;
    while (x < w - 4) x += 4;
    return x;
}

inline int reflect101(int i, int n)
{
    if (i < 0) i = -i;
    if (i >= n) i = 2 * n - 2 - i;
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

struct LevelPlan {
    int w, h, pitch;
    size_t img_off;
    int tile_h;                 // level kernel tile height (32 or 16)
    int tiles_x, tiles_y, ncells;
    int nCols, nRows, wCell, hCell, maxBorderX, maxBorderY;
    int nIni;
    float hX;
    int root_x[ORB_MAX_ROOTS + 1];
    int key_cap, list_cap;
    size_t key_off, ret_off, cell_off;
};

struct Plan {
    int W = 0, H = 0;
    LevelPlan lv[ORB_MAX_LEVELS];
    size_t img_bytes = 0, nkeys = 0, nret = 0, ncells = 0;
    int out_cap = 0;
};

}  // namespace orb_detail
using namespace orb_detail;

struct orb_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // Owners: the creator's handle plus every extractor, batch, server,
    // vocabulary and rectifier made on it; the last release frees it, so
    // destroying the handle before its users is safe.
    std::atomic<int> refs{1};
    // windowed searches: rounds before the sequential pass, and what the
    // last search used
    int round_limit = 48;
    int last_rounds = 0, last_sequential = 0;
    // brute-force matcher scratch
    uint8_t* d_q = nullptr;
    uint8_t* d_t = nullptr;
    int* d_out = nullptr;
    size_t q_cap = 0, t_cap = 0;
    // dense k=2 search scratch
    uint2* d_part = nullptr;
    size_t part_cap = 0;
    uint8_t* d_kq = nullptr;
    uint8_t* d_km = nullptr;
    int* d_kout = nullptr;
    size_t kq_cap = 0, km_cap = 0;
    // colour input staging
    uint8_t* d_color = nullptr;
    size_t color_cap = 0;
    // windowed matcher: inputs + scratch, candidate lists
    void* d_win = nullptr;
    size_t win_cap = 0;
    void* d_wcand = nullptr;
    size_t wcand_cap = 0;
    // map-point scratch: observation descriptors, offsets, chosen rows
    uint8_t* d_mp = nullptr;
    int* d_mpo = nullptr;
    int* d_mpb = nullptr;
    uint8_t* d_mpd = nullptr;
    size_t mp_cap = 0, mpo_cap = 0, mpb_cap = 0, mpd_cap = 0;
    // Frame keypoint ops: keypoints in and out, the RGB-D depth image
    void* d_fr = nullptr;
    size_t fr_cap = 0;

    int grow(void** p, size_t* cap, size_t bytes)
    {
        if (bytes <= *cap) return ORB_OK;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            return ORB_E_OOM;
        }
        *cap = bytes;
        return ORB_OK;
    }
};

// Releases one owner of a context; the last frees its buffers and stream.
void orb_context_release_ref(orb_context* c);
inline void orb_context_retain_ref(orb_context* c)
{
    if (c) c->refs.fetch_add(1, std::memory_order_relaxed);
}

// A context pointer that owns one reference while set.
struct CtxRef {
    orb_context* p = nullptr;
    CtxRef() = default;
    CtxRef(const CtxRef&) = delete;
    CtxRef& operator=(const CtxRef&) = delete;
    CtxRef& operator=(orb_context* c)
    {
        if (c != p) {
            orb_context_retain_ref(c);
            orb_context_release_ref(p);
            p = c;
        }
        return *this;
    }
    ~CtxRef() { orb_context_release_ref(p); }
    operator orb_context*() const { return p; }
    orb_context* operator->() const { return p; }
};

struct orb_extractor {
    CtxRef ctx;
    int nfeatures = 0, nlevels = 0, iniTh = 0, minTh = 0;
    float scaleFactorF = 1.2f;
    double scaleFactor = 1.2;
    // 1: resize and Gaussian 8U rounded as OpenCV's SSE2 kernels do (every
    // x86 build of OpenCV 2.4/3.x); 0: as its generic scalar code
    int rounding = 1;
    // level kernel tile height per level (64 x 32 or 64 x 16 tiles)
    int tile_h[ORB_MAX_LEVELS] = {32, 32, 32, 32, 32, 32, 32, 32, 32, 32, 32, 32, 32, 32, 32, 32};
    std::vector<float> scale, invScale, sigma2, invSigma2;
    std::vector<int> quota;

    Plan plan;
    unsigned gen = 0;             // bumped whenever buffers or plan change
    char* d_mem = nullptr;
    size_t mem_bytes = 0;
    uint8_t* d_stage = nullptr;   // contiguous landing buffer for host frames
    ImageJob hjob{};
    ImageJob* d_job = nullptr;
    RunTable* h_rt = nullptr;     // pinned
    RunTable* d_rt = nullptr;
    // stereo state (this extractor as the left view)
    float* d_uR = nullptr;
    float* d_depth = nullptr;
    int* d_msad = nullptr;        // per left keypoint: SAD of its stereo match
    int* d_misc = nullptr;        // [1] nl, [2] nr of caller-supplied stereo inputs
    int4* d_srec = nullptr;       // right keypoints bucketed by row (stereo scratch) ...
    uint8_t* d_sdesc = nullptr;   // ... and their descriptors
    int2* d_rows = nullptr;       // per-row candidate range in the bucketed right keypoints
    OrbKeyPoint* d_kin = nullptr; // uploaded keypoints: left then right
    uint8_t* d_din = nullptr;
    int in_cap = 0;
    // two-view run (this = left, pair_right = right)
    const orb_extractor* pair_right = nullptr;
    unsigned pair_gen_l = 0, pair_gen_r = 0;
    ImageJob* d_pair_jobs = nullptr;
    RunTable* h_pair_rt = nullptr;
    RunTable* d_pair_rt = nullptr;
    // device copies of the stereo arguments: [0] extractor-resident keypoints,
    // [1] caller-uploaded keypoints
    StereoArgs* d_sargs = nullptr;
    StereoArgs h_sargs[2];
    bool sargs_valid[2] = {false, false};

    int upload_stereo_args(int which, const StereoArgs& a)
    {
        if (!d_sargs && hipMalloc(&d_sargs, 2 * sizeof(StereoArgs)) != hipSuccess) {
            d_sargs = nullptr;
            return ORB_E_OOM;
        }
        if (sargs_valid[which] && std::memcmp(&a, &h_sargs[which], sizeof(a)) == 0) return ORB_OK;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(d_sargs + which, &a, sizeof(a), hipMemcpyHostToDevice));
        h_sargs[which] = a;
        sargs_valid[which] = true;
        return ORB_OK;
    }

    // ORBextractor::ORBextractor (ORBextractor.cc:469-526): scale tables in
    // float (mvScaleFactor[i-1]*scaleFactor with the double member), inverse
    // tables, and the geometric per-level quota with the remainder on top.
    void init_tables()
    {
        scale.assign(nlevels, 1.f);
        sigma2.assign(nlevels, 1.f);
        for (int i = 1; i < nlevels; i++) {
            scale[i] = (float)(scale[i - 1] * scaleFactor);
            sigma2[i] = scale[i] * scale[i];
        }
        invScale.resize(nlevels);
        invSigma2.resize(nlevels);
        for (int i = 0; i < nlevels; i++) {
            invScale[i] = 1.0f / scale[i];
            invSigma2[i] = 1.0f / sigma2[i];
        }
        quota.assign(nlevels, 0);
        const float factor = (float)(1.0f / scaleFactor);
        float nDesired = nfeatures * (1 - factor) / (1 - (float)std::pow((double)factor, (double)nlevels));
        int sum = 0;
        for (int l = 0; l < nlevels - 1; l++) {
            quota[l] = cv_round(nDesired);
            sum += quota[l];
            nDesired *= factor;
        }
        quota[nlevels - 1] = std::max(nfeatures - sum, 0);
    }

    // Level sizes (ComputePyramid, :1663), FAST cell grids (:1047-1071) and
    // quadtree roots (DistributeOctTree, :695-720) for a W x H image.
    int make_plan(int W, int H, Plan& p) const
    {
        p = Plan();
        p.W = W;
        p.H = H;
        size_t img = 0, nk = 0, nr = 0, ncell = 0;
        int out = 0;
        for (int l = 0; l < nlevels; l++) {
            LevelPlan& L = p.lv[l];
            L.w = cv_round((float)W * invScale[l]);
            L.h = cv_round((float)H * invScale[l]);
            if (L.w < 40 || L.h < 40) return ORB_E_ARG;
            L.pitch = (int)align_up((size_t)L.w, 64);
            L.img_off = img;
            img += align_up((size_t)L.pitch * L.h, 256);
            L.tile_h = tile_h[l];
            L.tiles_x = (L.w + ORB_TILE_W - 1) / ORB_TILE_W;
            L.tiles_y = (L.h + L.tile_h - 1) / L.tile_h;
            const int minBorderX = ORB_MIN_BORDER, minBorderY = ORB_MIN_BORDER;
            L.maxBorderX = L.w - ORB_EDGE + 3;
            L.maxBorderY = L.h - ORB_EDGE + 3;
            const float width = (float)(L.maxBorderX - minBorderX);
            const float height = (float)(L.maxBorderY - minBorderY);
            L.nCols = (int)(width / 30.f);
            L.nRows = (int)(height / 30.f);
            if (L.nCols <= 0 || L.nRows <= 0 || L.nCols > 255 || L.nRows > 255) return ORB_E_ARG;
            L.wCell = (int)std::ceil(width / L.nCols);
            L.hCell = (int)std::ceil(height / L.nRows);
            // A keypoint's offset from its cell's origin is < cell + 3 and is
            // packed in 6 bits; with nCols = (int)(width / 30) a cell is at
            // most 59 px (an integer width or height in [30, 60), one cell).
            if (L.wCell > 60 || L.hCell > 60) return ORB_E_ARG;
            L.ncells = L.nCols * L.nRows;
            L.nIni = (int)std::round((float)(L.maxBorderX - minBorderX) / (L.maxBorderY - minBorderY));
            if (L.nIni <= 0 || L.nIni > ORB_MAX_ROOTS) return ORB_E_ARG;
            L.hX = (float)(L.maxBorderX - minBorderX) / L.nIni;
            for (int i = 0; i <= L.nIni; i++) L.root_x[i] = (int)(L.hX * (float)i);
            // Strict 3x3 non-max suppression keeps at most one pixel per 2x2
            // block of a cell; candidates of both thresholds are stored.
            L.key_cap = 2 * L.ncells * ((L.wCell + 1) / 2) * ((L.hCell + 1) / 2);
            L.cell_off = ncell;
            ncell += (size_t)L.ncells;
            L.list_cap = std::max(quota[l] + 3, 4 * L.nIni);
            if (L.list_cap > ORB_LIST_CAP) return ORB_E_ARG;
            L.key_off = nk;
            nk += (size_t)L.key_cap;
            L.ret_off = nr;
            nr += (size_t)L.list_cap;
            out += L.list_cap;
        }
        p.img_bytes = img;
        p.nkeys = nk;
        p.nret = nr;
        p.ncells = ncell;
        p.out_cap = out;
        return ORB_OK;
    }

    // Device buffers and the job descriptor for the current image size.
    int ensure(int W, int H)
    {
        if (plan.W == W && plan.H == H && d_mem) return ORB_OK;
        Plan p;
        int rc = make_plan(W, H, p);
        if (rc) return rc;
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t o = off;
            off = align_up(off + bytes, 256);
            return o;
        };
        const size_t o_stage = take((size_t)W * H);   // host frames land here (one linear copy)
        const size_t o_lvl = take(p.img_bytes), o_blur = take(p.img_bytes), o_cell = take(p.ncells);
        const size_t o_keys = take(p.nkeys * 8), o_keys2 = take(p.nkeys * 8), o_node = take(p.nkeys * 4);
        const size_t o_ret = take(p.nret * 8);
        const size_t o_ctr = take(ORB_CNT_SLOTS * 4);
        const size_t o_kps = take((size_t)p.out_cap * sizeof(OrbKeyPoint));
        const size_t o_desc = take((size_t)p.out_cap * 32);
        const size_t o_uR = take((size_t)p.out_cap * 4), o_dep = take((size_t)p.out_cap * 4);
        const size_t o_msad = take((size_t)p.out_cap * 4);
        const size_t o_misc = take(64);
        const size_t o_srec = take((size_t)p.out_cap * 16), o_sdesc = take((size_t)p.out_cap * 32);
        const size_t o_rows = take((size_t)nlevels * ((size_t)H + 2) * 8);
        size_t ntab = 0;
        for (int l = 1; l < nlevels; l++)
            // tile-row windows sized for the smallest tile height, so that a
            // change of tile heights never reallocates (the frames stay put)
            ntab += (size_t)p.lv[l].w + p.lv[l].h + p.lv[l].tiles_x + (p.lv[l].h + 15) / 16;
        const size_t o_tab = take(ntab * 8);
        const size_t total = off;
        if (total > mem_bytes) {
            if (d_mem) (void)hipFree(d_mem);
            d_mem = nullptr;
            mem_bytes = 0;
            if (hipMalloc(&d_mem, total) != hipSuccess) {
                d_mem = nullptr;
                return ORB_E_OOM;
            }
            mem_bytes = total;
        }
        plan = p;
        gen++;
        ImageJob& J = hjob;
        std::memset(&J, 0, sizeof(J));
        // The input frame is written straight into level 0 (64-byte aligned
        // rows), which the level-0 kernel then reads in place.
        J.src = (const uint8_t*)(d_mem + o_lvl + p.lv[0].img_off);
        J.src_pitch = p.lv[0].pitch;
        d_stage = (uint8_t*)(d_mem + o_stage);
        J.nlevels = nlevels;
        J.iniTh = iniTh;
        J.minTh = minTh;
        J.out_cap = p.out_cap;
        J.counters = (int*)(d_mem + o_ctr);
        J.kps = (OrbKeyPoint*)(d_mem + o_kps);
        J.desc = (uint8_t*)(d_mem + o_desc);
        for (int l = 0; l < nlevels; l++) {
            const LevelPlan& P = p.lv[l];
            LevelJob& L = J.lv[l];
            L.w = P.w;
            L.h = P.h;
            L.pitch = P.pitch;
            L.tiles_x = P.tiles_x;
            L.tiles_y = P.tiles_y;
            L.nCols = P.nCols;
            L.nRows = P.nRows;
            L.wCell = P.wCell;
            L.hCell = P.hCell;
            // exact for x * cell < 2^32 (x < 2^16, cell <= 60)
            L.wrcp = (unsigned)(((1ull << 32) + (unsigned)P.wCell - 1) / (unsigned)P.wCell);
            L.hrcp = (unsigned)(((1ull << 32) + (unsigned)P.hCell - 1) / (unsigned)P.hCell);
            L.maxBorderX = P.maxBorderX;
            L.maxBorderY = P.maxBorderY;
            L.quota = quota[l];
            L.nIni = P.nIni;
            L.hX = P.hX;
            L.scale = scale[l];
            for (int i = 0; i <= P.nIni; i++) L.root_x[i] = P.root_x[i];
            L.key_cap = P.key_cap;
            L.ncells = P.ncells;
            L.list_cap = P.list_cap;
            L.rs_end = rounding ? sse2_resize_end(P.w) : 0;
            L.bs_end = rounding ? (P.w & ~3) : 0;
            L.lvl = (uint8_t*)(d_mem + o_lvl + P.img_off);
            L.blur = (uint8_t*)(d_mem + o_blur + P.img_off);
            L.cellflag = (uint8_t*)(d_mem + o_cell + P.cell_off);
            L.keys = (unsigned long long*)(d_mem + o_keys) + P.key_off;
            L.keys2 = (unsigned long long*)(d_mem + o_keys2) + P.key_off;
            L.node = (int*)(d_mem + o_node) + P.key_off;
            L.ret = (unsigned long long*)(d_mem + o_ret) + P.ret_off;
        }
        // cv::resize INTER_LINEAR taps of every level from the level above
        // (ComputePyramid, ORBextractor.cc:1677): per destination column and
        // row, the source index and the two 11-bit weights, with the float
        // statements of OpenCV's generic path.
        {
            std::vector<int2> tab(ntab);
            size_t t = 0;
            for (int l = 1; l < nlevels; l++) {
                const int w = p.lv[l].w, h = p.lv[l].h, sw = p.lv[l - 1].w, sh = p.lv[l - 1].h;
                J.lv[l].xtab = (const int2*)(d_mem + o_tab) + t;
                const double scale_x = 1. / ((double)w / sw);
                const int rs_end = J.lv[l].rs_end;
                for (int x = 0; x < w; x++, t++) {
                    float fx = (float)((x + 0.5) * scale_x - 0.5);
                    int sx = (int)std::floor(fx);
                    fx -= sx;
                    if (sx < 0) {
                        fx = 0.f;
                        sx = 0;
                    }
                    if (sx >= sw - 1) {
                        fx = 0.f;
                        sx = sw - 1;
                    }
                    tab[t].x = sx | ((x < rs_end ? 1 : 0) << 30);
                    tab[t].y = (int)std::nearbyint((1.f - fx) * 2048.f) | ((int)std::nearbyint(fx * 2048.f) << 16);
                }
                const size_t tx = t - w;
                // ytab == xtab + w: the level kernel reads both through xtab.
                J.lv[l].ytab = (const int2*)(d_mem + o_tab) + t;
                const double scale_y = 1. / ((double)h / sh);
                for (int y = 0; y < h; y++, t++) {
                    float fy = (float)((y + 0.5) * scale_y - 0.5);
                    const int sy = (int)std::floor(fy);
                    fy -= sy;
                    // the two source rows, clamped (the vertical taps are not
                    // zeroed at the border): sy | sy + 1 << 16
                    tab[t].x = std::min(std::max(sy, 0), sh - 1) | (std::min(std::max(sy + 1, 0), sh - 1) << 16);
                    tab[t].y = (int)std::nearbyint((1.f - fy) * 2048.f) | ((int)std::nearbyint(fy * 2048.f) << 16);
                }
                const size_t ty = t - h;
                // Per tile column and row: the source range its taps read,
                // halo (REFLECT_101) included, as the level kernel stages it.
                J.lv[l].xwin = (const int2*)(d_mem + o_tab) + t;
                for (int c = 0; c < p.lv[l].tiles_x; c++, t++) {
                    int lo = INT_MAX, hi = -1;
                    for (int i = 0; i < ORB_TILE_W + 8; i++) {
                        const int sx = tab[tx + reflect101(c * ORB_TILE_W - 4 + i, w)].x & 0x3FFFFFFF;
                        lo = std::min(lo, sx);
                        hi = std::max(hi, sx + 1 < sw ? sx + 1 : sx);
                    }
                    tab[t].x = lo;
                    tab[t].y = hi;
                }
                J.lv[l].ywin = (const int2*)(d_mem + o_tab) + t;
                for (int r = 0; r < p.lv[l].tiles_y; r++, t++) {
                    int lo = INT_MAX, hi = -1;
                    const int th = p.lv[l].tile_h;
                    for (int i = 0; i < th + 8; i++) {
                        const int sy01 = tab[ty + reflect101(r * th - 4 + i, h)].x;
                        lo = std::min(lo, sy01 & 0xFFFF);
                        hi = std::max(hi, sy01 >> 16);
                    }
                    tab[t].x = lo;
                    tab[t].y = hi;
                }
            }
            if (ntab) HIP_TRY(hipMemcpy(d_mem + o_tab, tab.data(), ntab * 8, hipMemcpyHostToDevice));
        }
        d_uR = (float*)(d_mem + o_uR);
        d_depth = (float*)(d_mem + o_dep);
        d_msad = (int*)(d_mem + o_msad);
        d_misc = (int*)(d_mem + o_misc);
        d_srec = (int4*)(d_mem + o_srec);
        d_sdesc = (uint8_t*)(d_mem + o_sdesc);
        d_rows = (int2*)(d_mem + o_rows);
        if (!d_job && hipMalloc(&d_job, sizeof(ImageJob)) != hipSuccess) return ORB_E_OOM;
        if (!d_rt && hipMalloc(&d_rt, sizeof(RunTable)) != hipSuccess) return ORB_E_OOM;
        if (!h_rt && hipHostMalloc(&h_rt, sizeof(RunTable)) != hipSuccess) return ORB_E_OOM;
        const orb_extractor* one[1] = {this};
        int rc_t = build_table(one, 1, *h_rt);
        if (rc_t) return rc_t;
        HIP_TRY(hipMemcpy(d_job, &hjob, sizeof(ImageJob), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_rt, h_rt, sizeof(RunTable), hipMemcpyHostToDevice));
        return ORB_OK;
    }

    // Launch table of a run over these jobs.  The level kernel's tile height
    // is a template parameter, so every job of a run must use the same one at
    // each level (ORB_E_ARG otherwise).
    static int build_table(const orb_extractor* const* exs, int n, RunTable& rt)
    {
        std::memset(&rt, 0, sizeof(rt));
        rt.njobs = n;
        int maxl = 0;
        for (int j = 0; j < n; j++) maxl = std::max(maxl, exs[j]->nlevels);
        rt.max_levels = maxl;
        for (int l = 0; l < maxl; l++) {
            rt.tile_h[l] = 0;
            for (int j = 0; j < n; j++) {
                if (l >= exs[j]->nlevels) continue;
                if (rt.tile_h[l] == 0) rt.tile_h[l] = exs[j]->plan.lv[l].tile_h;
                else if (rt.tile_h[l] != exs[j]->plan.lv[l].tile_h) return ORB_E_ARG;
            }
        }
        for (int l = 0; l < maxl; l++) {
            int acc = 0;
            for (int j = 0; j < n; j++) {
                rt.tile_start[l][j] = acc;
                if (l < exs[j]->nlevels) acc += exs[j]->plan.lv[l].tiles_x * exs[j]->plan.lv[l].tiles_y;
            }
            rt.tile_start[l][n] = acc;
            // equal tile counts: the kernels find a block's job by one division
            bool eq = n > 0;
            for (int j = 0; j < n; j++)
                eq = eq && rt.tile_start[l][j + 1] - rt.tile_start[l][j] == rt.tile_start[l][1] - rt.tile_start[l][0];
            rt.tiles_per_job[l] = eq ? rt.tile_start[l][1] - rt.tile_start[l][0] : 0;
        }
        int acc = 0;
        for (int j = 0; j < n; j++)
            for (int l = 0; l < maxl; l++) {
                rt.desc_start[j * maxl + l] = acc;
                if (l < exs[j]->nlevels) acc += (exs[j]->plan.lv[l].list_cap + ORB_DESC_KPB - 1) / ORB_DESC_KPB;
            }
        rt.desc_start[n * maxl] = acc;
        bool eq = n > 0;
        for (int j = 0; j < n; j++)
            eq = eq && rt.desc_start[(j + 1) * maxl] - rt.desc_start[j * maxl] == rt.desc_start[maxl];
        rt.desc_per_job = eq ? rt.desc_start[maxl] : 0;
        return ORB_OK;
    }

    // Input frame into level 0 of the pyramid (hjob.src, src_pitch bytes a
    // row), where the level-0 kernel reads it in place.  A packed host frame
    // lands in the staging buffer with one linear copy (a pitched host copy
    // is issued row by row and is far slower) and is then laid out on the
    // device.
    int upload_image(const uint8_t* img, size_t stride, hipMemcpyKind kind, hipStream_t st = nullptr)
    {
        if (!st) st = ctx->stream;
        uint8_t* dst = const_cast<uint8_t*>(hjob.src);
        const size_t dp = (size_t)hjob.src_pitch, W = (size_t)plan.W;
        if (kind == hipMemcpyHostToDevice && stride == W) {
            HIP_TRY(hipMemcpyAsync(d_stage, img, W * plan.H, kind, st));
            img = d_stage;
            kind = hipMemcpyDeviceToDevice;
        }
        if (stride == dp)
            HIP_TRY(hipMemcpyAsync(dst, img, dp * plan.H, kind, st));
        else
            HIP_TRY(hipMemcpy2DAsync(dst, dp, img, stride, W, plan.H, kind, st));
        return ORB_OK;
    }

    int check_errors()
    {
        int c[ORB_CNT_SLOTS];
        HIP_TRY(hipMemcpy(c, hjob.counters, sizeof(c), hipMemcpyDeviceToHost));
        return c[ORB_CNT_ERR] ? ORB_E_OVERFLOW : ORB_OK;
    }

    int read_results(orb_keypoint* kps, int cap, uint8_t* desc, int* n)
    {
        int c[ORB_CNT_SLOTS];
        HIP_TRY(hipMemcpyAsync(c, hjob.counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (c[ORB_CNT_ERR]) return ORB_E_OVERFLOW;
        const int cnt = c[ORB_CNT_NOUT];
        *n = cnt;
        if (cnt > cap) return ORB_E_CAPACITY;
        if (cnt > 0) {
            if (kps)
                HIP_TRY(hipMemcpyAsync(kps, hjob.kps, (size_t)cnt * sizeof(orb_keypoint), hipMemcpyDeviceToHost,
                                       ctx->stream));
            if (desc)
                HIP_TRY(hipMemcpyAsync(desc, hjob.desc, (size_t)cnt * 32, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
        return ORB_OK;
    }

    // Two-view run table (left = this, right = r) for the stereo frame.
    int ensure_pair(const orb_extractor* r)
    {
        if (pair_right == r && pair_gen_l == gen && pair_gen_r == r->gen) return ORB_OK;
        if (!d_pair_jobs && hipMalloc(&d_pair_jobs, 2 * sizeof(ImageJob)) != hipSuccess) return ORB_E_OOM;
        if (!d_pair_rt && hipMalloc(&d_pair_rt, sizeof(RunTable)) != hipSuccess) return ORB_E_OOM;
        if (!h_pair_rt && hipHostMalloc(&h_pair_rt, sizeof(RunTable)) != hipSuccess) return ORB_E_OOM;
        const orb_extractor* two[2] = {this, r};
        const int rc_t = build_table(two, 2, *h_pair_rt);
        if (rc_t) return rc_t;
        ImageJob jobs[2] = {hjob, r->hjob};
        HIP_TRY(hipMemcpy(d_pair_jobs, jobs, sizeof(jobs), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_pair_rt, h_pair_rt, sizeof(RunTable), hipMemcpyHostToDevice));
        pair_right = r;
        pair_gen_l = gen;
        pair_gen_r = r->gen;
        return ORB_OK;
    }

    int ensure_inputs(int cap)
    {
        if (cap <= in_cap) return ORB_OK;
        if (d_kin) (void)hipFree(d_kin);
        if (d_din) (void)hipFree(d_din);
        d_kin = nullptr;
        d_din = nullptr;
        in_cap = 0;
        if (hipMalloc(&d_kin, 2 * (size_t)cap * sizeof(OrbKeyPoint)) != hipSuccess) return ORB_E_OOM;
        if (hipMalloc(&d_din, 2 * (size_t)cap * 32) != hipSuccess) return ORB_E_OOM;
        in_cap = cap;
        return ORB_OK;
    }

    void stereo_args(StereoArgs& a, const orb_extractor* r, const ImageJob* dl_job, const ImageJob* dr_job,
                     float bf, float b) const
    {
        std::memset(&a, 0, sizeof(a));
        a.left = dl_job;
        a.right = dr_job;
        a.nrows = plan.lv[0].h;
        a.bf = bf;
        a.mb = b;
        for (int l = 0; l < nlevels; l++) {
            a.scale[l] = scale[l];
            a.inv_scale[l] = invScale[l];
        }
        a.uRight = d_uR;
        a.depth = d_depth;
        a.match_sad = d_msad;
        a.srec = d_srec;
        a.sdesc = d_sdesc;
        a.rows = d_rows;
        a.cap_r = std::min(plan.out_cap, r->plan.out_cap);
        a.nlev = nlevels;
    }

    ~orb_extractor()
    {
        if (d_mem) (void)hipFree(d_mem);
        if (d_job) (void)hipFree(d_job);
        if (d_rt) (void)hipFree(d_rt);
        if (h_rt) (void)hipHostFree(h_rt);
        if (d_kin) (void)hipFree(d_kin);
        if (d_din) (void)hipFree(d_din);
        if (d_pair_jobs) (void)hipFree(d_pair_jobs);
        if (d_pair_rt) (void)hipFree(d_pair_rt);
        if (h_pair_rt) (void)hipHostFree(h_pair_rt);
        if (d_sargs) (void)hipFree(d_sargs);
    }
};

