// orb_winmatch.cpp -- C ABI of ORBmatcher::SearchForInitialization over the
// Frame keypoint grid, for many frame pairs per call (kernels in
// orb_winmatch.hip).  Two launches with one host read between them: the
// candidate counts size the (i2, distance) lists exactly.  Device buffers
// are the context's, grown on demand and kept between calls.
#include <chrono>
#include <cstring>
#include <vector>

#include "orb_internal.h"

using namespace orb_detail;

extern "C" hipError_t orb_launch_win_count(const WinPair* d_pairs, int npairs, int max_n1, hipStream_t st);
extern "C" hipError_t orb_launch_win_scan(const WinPair* d_pairs, int npairs, int* d_totals, hipStream_t st);
extern "C" hipError_t orb_launch_win_fill(const WinPair* d_pairs, int npairs, int max_n1, hipStream_t st);
extern "C" hipError_t orb_launch_win_rounds(const WinPair* d_pairs, int npairs, int max_n1, int k0, int k1,
                                            hipStream_t st);
extern "C" hipError_t orb_launch_win_finish(const WinPair* d_pairs, int npairs, int sequential, hipStream_t st);

namespace {

size_t up(size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

namespace {

// steps > 0: after the call's own run, `steps` more runs of the device work
// with the inputs resident in HBM (vbPrevMatched restored before each), their
// mean wall time in *ms; the results returned are the call's.
int sfi_impl(orb_context* ctx, int n_pairs, const orb_keypoint* kp1, const uint8_t* desc1, const int32_t* off1,
             const orb_keypoint* kp2, const uint8_t* desc2, const int32_t* off2, const float* bounds,
             float* prev_matched, int window_size, float nn_ratio, int check_orientation, int32_t* matches12,
             int32_t* nmatches, int steps, float* ms)
{
    if (!ctx || n_pairs < 0 || (n_pairs && (!off1 || !off2 || !bounds || !matches12 || !nmatches))) return ORB_E_ARG;
    if (n_pairs == 0) return ORB_OK;
    const int N1 = off1[n_pairs], N2 = off2[n_pairs];
    if ((N1 && (!kp1 || !desc1 || !prev_matched)) || (N2 && (!kp2 || !desc2))) return ORB_E_ARG;
    int max_n1 = 0;
    for (int p = 0; p < n_pairs; p++) {
        const int a = off1[p + 1] - off1[p], b = off2[p + 1] - off2[p];
        if (a < 0 || b < 0 || off1[p] < 0 || off2[p] < 0) return ORB_E_ARG;
        if (a > 4096 || b > 4096) return ORB_E_CAPACITY;
        if (!(bounds[4 * p + 1] > bounds[4 * p]) || !(bounds[4 * p + 3] > bounds[4 * p + 2])) return ORB_E_ARG;
        max_n1 = std::max(max_n1, a);
    }
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    static_assert(sizeof(orb_keypoint) == sizeof(OrbKeyPoint), "keypoint layout");
    // inputs and per-pair scratch in one allocation
    const size_t s_kp1 = up((size_t)N1 * sizeof(OrbKeyPoint)), s_d1 = up((size_t)N1 * 32);
    const size_t s_kp2 = up((size_t)N2 * sizeof(OrbKeyPoint)), s_d2 = up((size_t)N2 * 32);
    const size_t s_prev = up((size_t)N1 * 8), s_sorted = up((size_t)N2 * 4);
    const size_t s_cells = up((size_t)n_pairs * (64 * 48 + 1) * 4), s_cnt = up((size_t)N1 * 4);
    const size_t s_off = up((size_t)(N1 + n_pairs) * 4), s_m12 = up((size_t)N1 * 4), s_nm = up((size_t)n_pairs * 4);
    const size_t s_pairs = up((size_t)n_pairs * sizeof(WinPair)), s_tot = up((size_t)n_pairs * 4);
    // round state: outcomes by parity, claims and claim counts by round mod 3, change flags
    const size_t s_out = up((size_t)N1 * 4), s_claim = up((size_t)N2 * ORB_WIN_SLOTS * 4), s_ccnt = up((size_t)N2 * 4);
    const size_t s_chg = up((size_t)n_pairs * ORB_WIN_MAXIT * 4);
    const size_t total = s_kp1 + s_d1 + s_kp2 + s_d2 + 2 * s_prev + s_sorted + s_cells + s_cnt + s_off + s_m12 +
                         s_nm + s_pairs + s_tot + 2 * s_out + 3 * s_claim + 3 * s_ccnt + s_chg;
    int rc = ctx->grow(&ctx->d_win, &ctx->win_cap, total);     // kept by the context between calls
    if (rc) return rc;
    char* m = (char*)ctx->d_win;
    size_t o = 0;
    auto take = [&](size_t s) {
        char* q = m + o;
        o += s;
        return q;
    };
    OrbKeyPoint* d_kp1 = (OrbKeyPoint*)take(s_kp1);
    uint8_t* d_d1 = (uint8_t*)take(s_d1);
    OrbKeyPoint* d_kp2 = (OrbKeyPoint*)take(s_kp2);
    uint8_t* d_d2 = (uint8_t*)take(s_d2);
    float2* d_prev = (float2*)take(s_prev);
    int* d_sorted = (int*)take(s_sorted);
    int* d_cells = (int*)take(s_cells);
    int* d_cnt = (int*)take(s_cnt);
    int* d_off = (int*)take(s_off);
    int* d_m12 = (int*)take(s_m12);
    int* d_nm = (int*)take(s_nm);
    WinPair* d_pairs = (WinPair*)take(s_pairs);
    int* d_tot = (int*)take(s_tot);
    int* d_out[2] = {(int*)take(s_out), (int*)take(s_out)};
    unsigned* d_claim[3] = {(unsigned*)take(s_claim), (unsigned*)take(s_claim), (unsigned*)take(s_claim)};
    int* d_ccnt[3] = {(int*)take(s_ccnt), (int*)take(s_ccnt), (int*)take(s_ccnt)};
    int* d_chg = (int*)take(s_chg);
    float2* d_prev0 = (float2*)take(s_prev);       // vbPrevMatched as passed, for timed repeats
    if (N1) {
        HIP_TRY(hipMemcpyAsync(d_kp1, kp1, (size_t)N1 * sizeof(OrbKeyPoint), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_d1, desc1, (size_t)N1 * 32, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_prev, prev_matched, (size_t)N1 * 8, hipMemcpyHostToDevice, st));
    }
    if (N2) {
        HIP_TRY(hipMemcpyAsync(d_kp2, kp2, (size_t)N2 * sizeof(OrbKeyPoint), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_d2, desc2, (size_t)N2 * 32, hipMemcpyHostToDevice, st));
    }
    std::vector<WinPair> pairs(n_pairs);
    for (int p = 0; p < n_pairs; p++) {
        WinPair& w = pairs[p];
        std::memset(&w, 0, sizeof(w));
        w.kp1 = d_kp1 + off1[p];
        w.d1 = d_d1 + (size_t)off1[p] * 32;
        w.kp2 = d_kp2 + off2[p];
        w.d2 = d_d2 + (size_t)off2[p] * 32;
        w.n1 = off1[p + 1] - off1[p];
        w.n2 = off2[p + 1] - off2[p];
        w.minX = bounds[4 * p];
        w.maxX = bounds[4 * p + 1];
        w.minY = bounds[4 * p + 2];
        w.maxY = bounds[4 * p + 3];
        w.prev = d_prev + off1[p];
        w.window = window_size;
        w.nnratio = nn_ratio;
        w.check_orientation = check_orientation ? 1 : 0;
        w.sorted = d_sorted + off2[p];
        w.cell_start = d_cells + (size_t)p * (64 * 48 + 1);
        w.cnt = d_cnt + off1[p];
        w.off = d_off + off1[p] + p;
        w.m12 = d_m12 + off1[p];
        w.nmatch = d_nm + p;
        for (int b = 0; b < 2; b++) w.out[b] = d_out[b] + off1[p];
        for (int b = 0; b < 3; b++) {
            w.claim[b] = d_claim[b] + (size_t)off2[p] * ORB_WIN_SLOTS;
            w.ccnt[b] = d_ccnt[b] + off2[p];
        }
        w.changed = d_chg + (size_t)p * ORB_WIN_MAXIT;
    }
    if (steps > 0 && N1) HIP_TRY(hipMemcpyAsync(d_prev0, d_prev, (size_t)N1 * 8, hipMemcpyDeviceToDevice, st));
    auto run_device = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(d_pairs, pairs.data(), n_pairs * sizeof(WinPair), hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_win_count(d_pairs, n_pairs, max_n1, st));
    HIP_TRY(orb_launch_win_scan(d_pairs, n_pairs, d_tot, st));
    std::vector<int> tot(n_pairs);
    HIP_TRY(hipMemcpyAsync(tot.data(), d_tot, n_pairs * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    size_t ncand = 0;
    std::vector<size_t> base(n_pairs);
    for (int p = 0; p < n_pairs; p++) {
        base[p] = ncand;
        ncand += (size_t)tot[p];
    }
    const int rg = ctx->grow(&ctx->d_wcand, &ctx->wcand_cap, std::max<size_t>(ncand, 1) * sizeof(int2));
    if (rg) return rg;
    for (int p = 0; p < n_pairs; p++) pairs[p].cand = (int2*)ctx->d_wcand + base[p];
    HIP_TRY(hipMemcpyAsync(d_pairs, pairs.data(), n_pairs * sizeof(WinPair), hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_win_fill(d_pairs, n_pairs, max_n1, st));
    // Rounds in chunks of 4 between reads of the change flags, until every
    // pair has settled; pairs still changing after ORB_WIN_MAXIT rounds get
    // the sequential pass.
    std::vector<int> chg((size_t)n_pairs * ORB_WIN_MAXIT);
    int k = 0;
    bool settled = false;
    const int limit = std::min(ctx->round_limit, ORB_WIN_MAXIT);
    while (!settled && k < limit) {
        const int k1 = std::min(limit, k + 4);
        HIP_TRY(orb_launch_win_rounds(d_pairs, n_pairs, max_n1, k, k1, st));
        k = k1;
        HIP_TRY(hipMemcpyAsync(chg.data(), d_chg, chg.size() * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        settled = true;
        for (int p = 0; p < n_pairs && settled; p++) {
            const int* c = chg.data() + (size_t)p * ORB_WIN_MAXIT;
            bool sp = false;
            for (int r = 1; r < k && !sp; r++) sp = c[r] == 0 && !(c[r - 1] & 2);
            settled = sp;
        }
    }
    ctx->last_rounds = k;
    ctx->last_sequential = settled ? 0 : 1;
    HIP_TRY(orb_launch_win_finish(d_pairs, n_pairs, settled ? 0 : 1, st));
    return ORB_OK;
    };
    rc = run_device();
    if (rc) return rc;
    if (steps > 0) {
        HIP_TRY(hipStreamSynchronize(st));
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < steps; i++) {
            if (N1) HIP_TRY(hipMemcpyAsync(d_prev, d_prev0, (size_t)N1 * 8, hipMemcpyDeviceToDevice, st));
            rc = run_device();
            if (rc) return rc;
        }
        HIP_TRY(hipStreamSynchronize(st));
        *ms = (float)(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps);
    }
    if (N1) {
        HIP_TRY(hipMemcpyAsync(matches12, d_m12, (size_t)N1 * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(prev_matched, d_prev, (size_t)N1 * 8, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipMemcpyAsync(nmatches, d_nm, n_pairs * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

}  // namespace

extern "C" {

int orb_search_for_initialization(orb_context* ctx, int n_pairs, const orb_keypoint* kp1, const uint8_t* desc1,
                                  const int32_t* off1, const orb_keypoint* kp2, const uint8_t* desc2,
                                  const int32_t* off2, const float* bounds, float* prev_matched, int window_size,
                                  float nn_ratio, int check_orientation, int32_t* matches12, int32_t* nmatches)
{
    return sfi_impl(ctx, n_pairs, kp1, desc1, off1, kp2, desc2, off2, bounds, prev_matched, window_size, nn_ratio,
                    check_orientation, matches12, nmatches, 0, nullptr);
}

int orb_search_for_initialization_time(orb_context* ctx, int n_pairs, const orb_keypoint* kp1, const uint8_t* desc1,
                                       const int32_t* off1, const orb_keypoint* kp2, const uint8_t* desc2,
                                       const int32_t* off2, const float* bounds, float* prev_matched, int window_size,
                                       float nn_ratio, int check_orientation, int32_t* matches12, int32_t* nmatches,
                                       int steps, float* ms)
{
    if (steps <= 0 || !ms) return ORB_E_ARG;
    return sfi_impl(ctx, n_pairs, kp1, desc1, off1, kp2, desc2, off2, bounds, prev_matched, window_size, nn_ratio,
                    check_orientation, matches12, nmatches, steps, ms);
}

}  // extern "C"

// ---------------------------------------------------------- SearchByProjection
extern "C" hipError_t orb_launch_proj_prep(ProjFrame* d_frames, int nframes, int max_nq, int rule, const float* d_pts,
                                           const int* d_pt_off, const float* d_scale, int nlevels, const float* d_cam,
                                           int mono, float th, hipStream_t st);
extern "C" hipError_t orb_launch_proj_count(const WinPair* d_grid, ProjFrame* d_frames, int nframes, int max_nq,
                                            int* d_totals, hipStream_t st);
extern "C" hipError_t orb_launch_proj_fill(ProjFrame* d_frames, int nframes, int max_nq, hipStream_t st);
extern "C" hipError_t orb_launch_proj_rounds(ProjFrame* d_frames, int nframes, int max_nq, int k0, int k1,
                                             hipStream_t st);
extern "C" hipError_t orb_launch_proj_finish(ProjFrame* d_frames, int nframes, int sequential, hipStream_t st);

namespace {

// Both overloads: rule 1 = SearchByProjection(F, vpMapPoints, th), rule 2 =
// SearchByProjection(Current, Last, th, bMono).  Point records are 24 bytes
// (orb_proj_point / orb_last_point).  Rounds run in chunks of 4 between
// reads of the per-round change flags; frames still changing after
// ORB_PROJ_MAXIT rounds get the sequential pass.
int proj_search(orb_context* ctx, int rule, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                const float* uright, const uint8_t* occupied, const int32_t* key_off, const float* bounds,
                const float* scale, int nlevels, const void* pts, const uint8_t* pdesc, const int32_t* pt_off,
                const float* cam, int mono, float th, float nnratio, int check_orientation, int32_t* pt_match,
                int32_t* key_point, int32_t* nmatches, int steps, float* ms)
{
    static_assert(sizeof(orb_proj_point) == 24 && sizeof(orb_last_point) == 24, "point records");
    if (!ctx || n_frames < 0 || nlevels < 1 || nlevels > ORB_MAX_LEVELS ||
        (n_frames && (!key_off || !pt_off || !bounds || !scale || !nmatches || (rule == 2 && !cam))))
        return ORB_E_ARG;
    if (n_frames == 0) return ORB_OK;
    const int N = key_off[n_frames], Q = pt_off[n_frames];
    if ((N && (!kp || !desc || !uright || !occupied || !key_point)) || (Q && (!pts || !pdesc || !pt_match)))
        return ORB_E_ARG;
    int max_nq = 0;
    for (int f = 0; f < n_frames; f++) {
        const int n = key_off[f + 1] - key_off[f], nq = pt_off[f + 1] - pt_off[f];
        if (n < 0 || nq < 0 || key_off[f] < 0 || pt_off[f] < 0) return ORB_E_ARG;
        if (n > 4096 || nq > (1 << 20)) return ORB_E_CAPACITY;
        if (!(bounds[4 * f + 1] > bounds[4 * f]) || !(bounds[4 * f + 3] > bounds[4 * f + 2])) return ORB_E_ARG;
        max_nq = std::max(max_nq, nq);
    }
    for (int i = 0; i < N; i++)
        if (kp[i].octave < 0 || kp[i].octave >= ORB_MAX_LEVELS) return ORB_E_ARG;
    // The rotation check bins (last angle - current angle) mod 360; keypoint
    // angles lie in [0, 360] (the reference asserts a bin in range, :1836).
    if (rule == 2 && check_orientation) {
        for (int i = 0; i < N; i++)
            if (!(kp[i].angle >= 0.f && kp[i].angle <= 360.f)) return ORB_E_ARG;
        const orb_last_point* lp = (const orb_last_point*)pts;
        for (int q = 0; q < Q; q++)
            if ((lp[q].flags & 1) && !(lp[q].angle >= 0.f && lp[q].angle <= 360.f)) return ORB_E_ARG;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int F = n_frames;
    const size_t s_kp = up((size_t)N * sizeof(OrbKeyPoint)), s_desc = up((size_t)N * 32), s_ur = up((size_t)N * 4),
                 s_occ = up((size_t)N), s_pts = up((size_t)Q * 24), s_pdesc = up((size_t)Q * 32),
                 s_ptoff = up((size_t)(F + 1) * 4), s_scale = up(ORB_MAX_LEVELS * 4), s_cam = up((size_t)F * 7 * 4),
                 s_sorted = up((size_t)N * 4), s_cells = up((size_t)F * (64 * 48 + 1) * 4),
                 s_q = up((size_t)Q * sizeof(ProjQuery)), s_cnt = up((size_t)Q * 4), s_off = up((size_t)(Q + F) * 4),
                 s_out = up((size_t)Q * 4), s_claim = up((size_t)N * 4), s_chg = up((size_t)F * ORB_PROJ_MAXIT * 4),
                 s_nm = up((size_t)F * 4), s_grid = up((size_t)F * sizeof(WinPair)),
                 s_frames = up((size_t)F * sizeof(ProjFrame)), s_tot = up((size_t)F * 4);
    const size_t total = s_kp + s_desc + s_ur + s_occ + s_pts + s_pdesc + s_ptoff + s_scale + s_cam + s_sorted +
                         s_cells + s_q + s_cnt + s_off + 3 * s_out + 4 * s_claim + s_chg + s_nm + s_grid + s_frames +
                         s_tot;
    int rc = ctx->grow(&ctx->d_win, &ctx->win_cap, total);
    if (rc) return rc;
    char* m = (char*)ctx->d_win;
    size_t o = 0;
    auto take = [&](size_t s) {
        char* q = m + o;
        o += s;
        return q;
    };
    OrbKeyPoint* d_kp = (OrbKeyPoint*)take(s_kp);
    uint8_t* d_desc = (uint8_t*)take(s_desc);
    float* d_ur = (float*)take(s_ur);
    uint8_t* d_occ = (uint8_t*)take(s_occ);
    float* d_pts = (float*)take(s_pts);
    uint8_t* d_pdesc = (uint8_t*)take(s_pdesc);
    int* d_ptoff = (int*)take(s_ptoff);
    float* d_scale = (float*)take(s_scale);
    float* d_cam = (float*)take(s_cam);
    int* d_sorted = (int*)take(s_sorted);
    int* d_cells = (int*)take(s_cells);
    ProjQuery* d_q = (ProjQuery*)take(s_q);
    int* d_cnt = (int*)take(s_cnt);
    int* d_off = (int*)take(s_off);
    int* d_out0 = (int*)take(s_out);
    int* d_out1 = (int*)take(s_out);
    int* d_ptm = (int*)take(s_out);
    int* d_claim0 = (int*)take(s_claim);
    int* d_claim1 = (int*)take(s_claim);
    int* d_claim2 = (int*)take(s_claim);
    int* d_keyp = (int*)take(s_claim);
    int* d_chg = (int*)take(s_chg);
    int* d_nm = (int*)take(s_nm);
    WinPair* d_grid = (WinPair*)take(s_grid);
    ProjFrame* d_frames = (ProjFrame*)take(s_frames);
    int* d_tot = (int*)take(s_tot);
    static_assert(sizeof(orb_keypoint) == sizeof(OrbKeyPoint), "keypoint layout");
    if (N) {
        HIP_TRY(hipMemcpyAsync(d_kp, kp, (size_t)N * sizeof(OrbKeyPoint), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_desc, desc, (size_t)N * 32, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_ur, uright, (size_t)N * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_occ, occupied, (size_t)N, hipMemcpyHostToDevice, st));
    }
    if (Q) {
        HIP_TRY(hipMemcpyAsync(d_pts, pts, (size_t)Q * 24, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_pdesc, pdesc, (size_t)Q * 32, hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipMemcpyAsync(d_ptoff, pt_off, (size_t)(F + 1) * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_scale, scale, (size_t)nlevels * 4, hipMemcpyHostToDevice, st));
    if (rule == 2) HIP_TRY(hipMemcpyAsync(d_cam, cam, (size_t)F * 7 * 4, hipMemcpyHostToDevice, st));
    std::vector<WinPair> grid(F);
    std::vector<ProjFrame> frames(F);
    for (int f = 0; f < F; f++) {
        const int k0 = key_off[f], q0 = pt_off[f];
        WinPair& w = grid[f];
        std::memset(&w, 0, sizeof(w));
        w.kp2 = d_kp + k0;
        w.n2 = key_off[f + 1] - k0;
        w.minX = bounds[4 * f];
        w.maxX = bounds[4 * f + 1];
        w.minY = bounds[4 * f + 2];
        w.maxY = bounds[4 * f + 3];
        w.sorted = d_sorted + k0;
        w.cell_start = d_cells + (size_t)f * (64 * 48 + 1);
        w.n1 = pt_off[f + 1] - q0;
        w.cnt = d_cnt + q0;
        w.off = d_off + q0 + f;
        ProjFrame& P = frames[f];
        std::memset(&P, 0, sizeof(P));
        P.kp = d_kp + k0;
        P.desc = d_desc + (size_t)k0 * 32;
        P.uright = d_ur + k0;
        P.occupied = d_occ + k0;
        P.n = w.n2;
        P.minX = w.minX;
        P.maxX = w.maxX;
        P.minY = w.minY;
        P.maxY = w.maxY;
        P.sorted = w.sorted;
        P.cell_start = w.cell_start;
        P.q = d_q + q0;
        P.qdesc = d_pdesc + (size_t)q0 * 32;
        P.nq = w.n1;
        P.cnt = w.cnt;
        P.off = w.off;
        P.out[0] = d_out0 + q0;
        P.out[1] = d_out1 + q0;
        P.claim[0] = d_claim0 + k0;
        P.claim[1] = d_claim1 + k0;
        P.claim[2] = d_claim2 + k0;
        P.changed = d_chg + (size_t)f * ORB_PROJ_MAXIT;
        P.rule = rule;
        P.nnratio = nnratio;
        P.check_orientation = check_orientation ? 1 : 0;
        P.pt_match = d_ptm + q0;
        P.key_point = d_keyp + k0;
        P.nmatch = d_nm + f;
    }
    auto run_device = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(d_grid, grid.data(), (size_t)F * sizeof(WinPair), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_frames, frames.data(), (size_t)F * sizeof(ProjFrame), hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_proj_prep(d_frames, F, max_nq, rule, d_pts, d_ptoff, d_scale, nlevels, d_cam, mono, th, st));
    HIP_TRY(orb_launch_proj_count(d_grid, d_frames, F, max_nq, d_tot, st));
    std::vector<int> tot(F);
    HIP_TRY(hipMemcpyAsync(tot.data(), d_tot, (size_t)F * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    size_t ncand = 0;
    for (int f = 0; f < F; f++) ncand += (size_t)tot[f];
    const int rg = ctx->grow(&ctx->d_wcand, &ctx->wcand_cap, std::max<size_t>(ncand, 1) * sizeof(unsigned));
    if (rg) return rg;
    size_t base = 0;
    for (int f = 0; f < F; f++) {
        frames[f].cand = (unsigned*)ctx->d_wcand + base;
        base += (size_t)tot[f];
    }
    HIP_TRY(hipMemcpyAsync(d_frames, frames.data(), (size_t)F * sizeof(ProjFrame), hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_proj_fill(d_frames, F, max_nq, st));
    // claim-resolution rounds until every frame has a round that changed nothing
    std::vector<int> chg((size_t)F * ORB_PROJ_MAXIT);
    int k = 0;
    bool settled = false;
    const int limit = std::min(ctx->round_limit, ORB_PROJ_MAXIT);
    while (!settled && k < limit) {
        const int k1 = std::min(limit, k + 4);
        HIP_TRY(orb_launch_proj_rounds(d_frames, F, max_nq, k, k1, st));
        k = k1;
        HIP_TRY(hipMemcpyAsync(chg.data(), d_chg, chg.size() * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        settled = true;
        for (int f = 0; f < F && settled; f++) {
            bool s = false;
            for (int r = 1; r < k && !s; r++) s = chg[(size_t)f * ORB_PROJ_MAXIT + r] == 0;
            settled = s;
        }
    }
    ctx->last_rounds = k;
    ctx->last_sequential = settled ? 0 : 1;
    HIP_TRY(orb_launch_proj_finish(d_frames, F, settled ? 0 : 1, st));
    return ORB_OK;
    };
    rc = run_device();
    if (rc) return rc;
    if (steps > 0) {        // repeats with the inputs resident in HBM (they are not modified)
        HIP_TRY(hipStreamSynchronize(st));
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < steps; i++) {
            rc = run_device();
            if (rc) return rc;
        }
        HIP_TRY(hipStreamSynchronize(st));
        *ms = (float)(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps);
    }
    if (Q) HIP_TRY(hipMemcpyAsync(pt_match, d_ptm, (size_t)Q * 4, hipMemcpyDeviceToHost, st));
    if (N) HIP_TRY(hipMemcpyAsync(key_point, d_keyp, (size_t)N * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(nmatches, d_nm, (size_t)F * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

}  // namespace

extern "C" {

int orb_context_set_round_limit(orb_context* ctx, int rounds)
{
    if (!ctx || rounds < 0 || rounds > ORB_WIN_MAXIT) return ORB_E_ARG;
    ctx->round_limit = rounds;
    return ORB_OK;
}

int orb_context_last_rounds(const orb_context* ctx, int* sequential)
{
    if (!ctx) return ORB_E_ARG;
    if (sequential) *sequential = ctx->last_sequential;
    return ctx->last_rounds;
}

int orb_search_by_projection(orb_context* ctx, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                             const float* uright, const uint8_t* occupied, const int32_t* key_off,
                             const float* bounds, const float* scale_factors, int nlevels,
                             const orb_proj_point* points, const uint8_t* point_desc, const int32_t* point_off,
                             float th, float nn_ratio, int32_t* point_match, int32_t* key_point, int32_t* nmatches)
{
    return proj_search(ctx, 1, n_frames, kp, desc, uright, occupied, key_off, bounds, scale_factors, nlevels, points,
                       point_desc, point_off, nullptr, 0, th, nn_ratio, 0, point_match, key_point, nmatches, 0,
                       nullptr);
}

int orb_search_by_projection_time(orb_context* ctx, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                                  const float* uright, const uint8_t* occupied, const int32_t* key_off,
                                  const float* bounds, const float* scale_factors, int nlevels,
                                  const orb_proj_point* points, const uint8_t* point_desc, const int32_t* point_off,
                                  float th, float nn_ratio, int32_t* point_match, int32_t* key_point,
                                  int32_t* nmatches, int steps, float* ms)
{
    if (steps <= 0 || !ms) return ORB_E_ARG;
    return proj_search(ctx, 1, n_frames, kp, desc, uright, occupied, key_off, bounds, scale_factors, nlevels, points,
                       point_desc, point_off, nullptr, 0, th, nn_ratio, 0, point_match, key_point, nmatches, steps,
                       ms);
}

int orb_search_by_projection_last(orb_context* ctx, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                                  const float* uright, const uint8_t* occupied, const int32_t* key_off,
                                  const float* bounds, const float* scale_factors, int nlevels, const float* camera,
                                  int mono, const orb_last_point* points, const uint8_t* point_desc,
                                  const int32_t* point_off, float th, int check_orientation, int32_t* point_match,
                                  int32_t* key_point, int32_t* nmatches)
{
    return proj_search(ctx, 2, n_frames, kp, desc, uright, occupied, key_off, bounds, scale_factors, nlevels, points,
                       point_desc, point_off, camera, mono, th, 0.f, check_orientation, point_match, key_point,
                       nmatches, 0, nullptr);
}

int orb_search_by_projection_last_time(orb_context* ctx, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                                       const float* uright, const uint8_t* occupied, const int32_t* key_off,
                                       const float* bounds, const float* scale_factors, int nlevels,
                                       const float* camera, int mono, const orb_last_point* points,
                                       const uint8_t* point_desc, const int32_t* point_off, float th,
                                       int check_orientation, int32_t* point_match, int32_t* key_point,
                                       int32_t* nmatches, int steps, float* ms)
{
    if (steps <= 0 || !ms) return ORB_E_ARG;
    return proj_search(ctx, 2, n_frames, kp, desc, uright, occupied, key_off, bounds, scale_factors, nlevels, points,
                       point_desc, point_off, camera, mono, th, 0.f, check_orientation, point_match, key_point,
                       nmatches, steps, ms);
}

}  // extern "C"
