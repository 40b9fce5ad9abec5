// orb_winmatch.cpp -- C ABI of ORBmatcher::SearchForInitialization over the
// Frame keypoint grid, for many frame pairs per call (kernels in
// orb_winmatch.hip).  Two launches with one host read between them: the
// candidate counts size the (i2, distance) lists exactly.  Device buffers
// are the context's, grown on demand and kept between calls.
#include <cstring>
#include <vector>

#include "orb_internal.h"

using namespace orb_detail;

extern "C" hipError_t orb_launch_win_count(const WinPair* d_pairs, int npairs, int max_n1, hipStream_t st);
extern "C" hipError_t orb_launch_win_scan(const WinPair* d_pairs, int npairs, int* d_totals, hipStream_t st);
extern "C" hipError_t orb_launch_win_match(const WinPair* d_pairs, int npairs, int max_n1, hipStream_t st);

namespace {

size_t up(size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

extern "C" {

int orb_search_for_initialization(orb_context* ctx, int n_pairs, const orb_keypoint* kp1, const uint8_t* desc1,
                                  const int32_t* off1, const orb_keypoint* kp2, const uint8_t* desc2,
                                  const int32_t* off2, const float* bounds, float* prev_matched, int window_size,
                                  float nn_ratio, int check_orientation, int32_t* matches12, int32_t* nmatches)
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
    const size_t total = s_kp1 + s_d1 + s_kp2 + s_d2 + s_prev + s_sorted + s_cells + s_cnt + s_off + s_m12 + s_nm +
                         s_pairs + s_tot;
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
    }
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
    rc = ctx->grow(&ctx->d_wcand, &ctx->wcand_cap, std::max<size_t>(ncand, 1) * sizeof(int2));
    if (rc) return rc;
    for (int p = 0; p < n_pairs; p++) pairs[p].cand = (int2*)ctx->d_wcand + base[p];
    HIP_TRY(hipMemcpyAsync(d_pairs, pairs.data(), n_pairs * sizeof(WinPair), hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_win_match(d_pairs, n_pairs, max_n1, st));
    if (N1) {
        HIP_TRY(hipMemcpyAsync(matches12, d_m12, (size_t)N1 * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(prev_matched, d_prev, (size_t)N1 * 8, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipMemcpyAsync(nmatches, d_nm, n_pairs * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

}  // extern "C"
