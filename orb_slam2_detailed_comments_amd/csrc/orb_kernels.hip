// orb_kernels.hip -- gfx950 kernels of the ORB front end.
//
// Pipeline per image (one launch per stage covers every job of a run):
//   level_kernel   (x nlevels) pyramid level l from level l-1 (8-bit fixed
//                  point bilinear, ORBextractor.cc:1677), its 7x7 sigma-2
//                  blur (ORBextractor.cc:1607-1615), FAST-9 with per-cell
//                  non-max suppression at both thresholds
//                  (ORBextractor.cc:1075-1142), from one LDS tile
//   octree_kernel  the per-cell iniThFAST/minThFAST retry, then
//                  DistributeOctTree as a per-level workgroup quadtree
//                  (ORBextractor.cc:688-1033)
//   describe_kernel IC_Angle + rotated BRIEF + output assembly
//                  (ORBextractor.cc:94-204, 1590-1648)
// plus stereo matching (Frame.cc:831-1082) and brute-force Hamming matching.
//
// All float code is compiled with -ffp-contract=off: every float statement
// that the reference evaluates is evaluated here with the same operations in
// the same order, so results are bit-identical to the CPU statement.
#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>

#include "orb_types.h"

__constant__ __align__(16) signed char c_pattern[1024] = {   // BRIEF pairs, x0,y0,x1,y1
#include "orb_pattern.inc"
};
__constant__ unsigned c_icmask[4 * 16 * 9];   // orientation disc byte masks [alignment][|v|][dword]
__constant__ int c_gauss[7];   // 7x7 sigma-2 integer taps (sum need not be 256)

// umax16: the orientation disc half-widths of the constructor
// (ORBextractor.cc:542-570).  Byte k of dword c of a patch row whose first
// dword starts `al` pixels left of u = -15 holds u = 4c + k - 15 - al; its
// mask byte is 0xFF iff |u| <= umax[|v|].
extern "C" hipError_t orb_upload_constants(const int* umax16, const int* gauss7)
{
    unsigned m[4 * 16 * 9];
    for (int al = 0; al < 4; al++)
        for (int av = 0; av < 16; av++)
            for (int c = 0; c < 9; c++) {
                unsigned w = 0;
                for (int k = 0; k < 4; k++) {
                    const int u = 4 * c + k - 15 - al;
                    if ((u < 0 ? -u : u) <= umax16[av]) w |= 0xFFu << (8 * k);
                }
                m[(al * 16 + av) * 9 + c] = w;
            }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_icmask), m, sizeof(m));
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_gauss), gauss7, 7 * sizeof(int));
}

// ------------------------------------------------------------------ helpers
// Loads of 16, 8 and 4 bytes at 4-byte aligned global addresses, marked as
// global so that they are issued as global (not flat) loads.
typedef unsigned orb_v4u __attribute__((ext_vector_type(4), aligned(4)));
typedef unsigned orb_v2u __attribute__((ext_vector_type(2), aligned(4)));
__device__ __forceinline__ uint4 gload16(const uint8_t* p)
{
    const orb_v4u v = *(const __attribute__((address_space(1))) orb_v4u*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 gload8(const uint8_t* p)
{
    const orb_v2u v = *(const __attribute__((address_space(1))) orb_v2u*)p;
    return make_uint2(v.x, v.y);
}
// 8 bytes that no thread writes during the kernel, at a wave-uniform address:
// constant address space, so that it is a scalar load.
__device__ __forceinline__ uint2 cload8(const void* p)
{
    const orb_v2u v = *(const __attribute__((address_space(4))) orb_v2u*)p;
    return make_uint2(v.x, v.y);
}
typedef unsigned orb_u32a1 __attribute__((aligned(1)));
__device__ __forceinline__ unsigned gload4u(const uint8_t* p)   // any alignment
{
    return *(const __attribute__((address_space(1))) orb_u32a1*)p;
}
// Stores through global (not flat) addresses.
__device__ __forceinline__ void gstore4(uint8_t* p, unsigned v) { *(__attribute__((address_space(1))) unsigned*)p = v; }
__device__ __forceinline__ void gstore1(uint8_t* p, uint8_t v) { *(__attribute__((address_space(1))) uint8_t*)p = v; }
__device__ __forceinline__ void gstore8(unsigned long long* p, unsigned long long v)
{
    *(__attribute__((address_space(1))) unsigned long long*)p = v;
}
__device__ __forceinline__ unsigned long long gload8u64(const unsigned long long* p)
{
    return *(const __attribute__((address_space(1))) unsigned long long*)p;
}
__device__ __forceinline__ uint8_t gload1(const uint8_t* p) { return *(const __attribute__((address_space(1))) uint8_t*)p; }
__device__ __forceinline__ void gstore2(uint8_t* p, unsigned short v)
{
    *(__attribute__((address_space(1))) unsigned short*)p = v;
}
__device__ __forceinline__ void gstore16(uint8_t* p, uint4 v)
{
    *(__attribute__((address_space(1))) orb_v4u*)p = orb_v4u{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ unsigned gload4(const uint8_t* p)
{
    return *(const __attribute__((address_space(1))) unsigned*)p;
}

__device__ __forceinline__ int reflect101(int i, int n)
{
    if (i < 0) i = -i;
    if (i >= n) i = 2 * n - 2 - i;
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

__device__ __forceinline__ int find_segment(const int* starts, int n, int x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Sum over the 16 lanes of a DPP row (all of them active); every lane of the
// row receives the total.
__device__ __forceinline__ unsigned row16_sum(unsigned v)
{
    v += (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
    v += (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
    v += (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false);   // row_half_mirror
    v += (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false);   // row_mirror
    return v;
}

// The dispatcher sends workgroup b to XCD b % 8, and each XCD has its own
// L2.  This bijection on [0, n) hands XCD x the contiguous logical range
// [x*q + min(x, r), ...) (q = n / 8, r = n % 8), so neighbouring tiles, whose
// halos share cache lines, and the keypoints of one image meet in one L2.
#define ORB_NUM_XCD 8
__device__ __forceinline__ int xcd_block(int b, int n)
{
    const int q = n / ORB_NUM_XCD, r = n % ORB_NUM_XCD;
    const int x = b % ORB_NUM_XCD, i = b / ORB_NUM_XCD;
    return x * q + min(x, r) + i;
}

__device__ __forceinline__ int lanes_below(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// ================================================================ level
// One workgroup per 64x32 output tile of pyramid level l:
//   1. level pixels for the tile plus a 4-px halo (REFLECT_101 at the level
//      edges) into LDS: copied for level 0, else cv::resize INTER_LINEAR 8U
//      from level l-1 (ComputePyramid, ORBextractor.cc:1677);
//   2. the level and its 7x7 sigma-2 blur (ORBextractor.cc:1607-1615);
//   3. FAST-9 arc strength for the tile plus a 1-px ring, computed exactly only
//      for pixels that pass a necessary cardinal-point test, after compaction;
//   4. per-cell non-max suppression at iniThFAST and at minThFAST: a neighbour
//      counts only if it lies in the same 30-px cell's detection area, which is
//      exactly cv::FAST run on that cell's window (ORBextractor.cc:1075-1142);
//   5. candidates of both thresholds appended to the level's key list, and a
//      per-cell flag set when the cell has an iniThFAST survivor.  The quadtree
//      kernel keeps minThFAST candidates only for cells without that flag,
//      which is the reference's per-cell retry.
#define LT_W ORB_TILE_W
#define LT_HALO 4
#define LP_W (LT_W + 2 * LT_HALO)      // 72 bytes = 18 dwords
#define LP_WD (LP_W / 4)
#define LK_CAP 1536                  // per-tile candidates, both thresholds

typedef unsigned short orb_us2 __attribute__((ext_vector_type(2)));

typedef short orb_s2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int fast_strength(const uint8_t* p, int stride)
{
    // Bresenham radius-3 circle in circular order (OpenCV makeOffsets(16)),
    // differences to the centre as 16-bit pairs D[j] = (d[2j], d[2j+1]),
    // each in [-255, 255].
    const unsigned v = p[0];
    const unsigned r[16] = {p[3 * stride],      p[3 * stride + 1],  p[2 * stride + 2],  p[stride + 3],
                            p[3],               p[-stride + 3],     p[-2 * stride + 2], p[-3 * stride + 1],
                            p[-3 * stride],     p[-3 * stride - 1], p[-2 * stride - 2], p[-stride - 3],
                            p[-3],              p[stride - 3],      p[2 * stride - 2],  p[3 * stride - 1]};
    const orb_s2 V = __builtin_bit_cast(orb_s2, v | (v << 16));
    orb_s2 D[8];
#pragma unroll
    for (int j = 0; j < 8; j++) D[j] = __builtin_bit_cast(orb_s2, r[2 * j] | (r[2 * j + 1] << 16)) - V;
    // Minimum and maximum over every run of 9 contiguous ring pixels, two
    // runs per packed op: pairs (k, k+1), then runs of 4, then 4 + 4 + 1.
    orb_s2 MN[8], MX[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const orb_s2 S = __builtin_bit_cast(orb_s2, __builtin_amdgcn_alignbit(__builtin_bit_cast(unsigned, D[(j + 1) & 7]),
                                                                            __builtin_bit_cast(unsigned, D[j]), 16));
        MN[j] = __builtin_elementwise_min(D[j], S);      // (d[2j], d[2j+1]) with (d[2j+1], d[2j+2])
        MX[j] = __builtin_elementwise_max(D[j], S);
    }
    orb_s2 M4[8], X4[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        M4[j] = __builtin_elementwise_min(MN[j], MN[(j + 1) & 7]);
        X4[j] = __builtin_elementwise_max(MX[j], MX[(j + 1) & 7]);
    }
    orb_s2 B = {-256, -256}, K = {256, 256};
#pragma unroll
    for (int j = 0; j < 8; j++) {
        B = __builtin_elementwise_max(B, __builtin_elementwise_min(__builtin_elementwise_min(M4[j], M4[(j + 2) & 7]),
                                                                   D[(j + 4) & 7]));
        K = __builtin_elementwise_min(K, __builtin_elementwise_max(__builtin_elementwise_max(X4[j], X4[(j + 2) & 7]),
                                                                   D[(j + 4) & 7]));
    }
    const int bright = max((int)B.x, (int)B.y), dark = min((int)K.x, (int)K.y);
    // A pixel is a FAST-9 corner at threshold t iff strength > t, and then
    // OpenCV's cornerScore equals strength - 1.
    const int s = max(bright, -dark);
    return s < 0 ? 0 : s;
}

// Necessary condition for strength > t, on two pixels at once (16-bit
// lanes): a 9-run covers two circularly adjacent cardinal points (ring
// positions 0, 4, 8, 12), all brighter than v+t or all darker than v-t.
__device__ __forceinline__ unsigned fast_maybe2(unsigned v, unsigned p0, unsigned p4, unsigned p8, unsigned p12,
                                                unsigned t2)
{
    const orb_us2 V = __builtin_bit_cast(orb_us2, v), T = __builtin_bit_cast(orb_us2, t2);
    const orb_us2 hi = V + T;                                   // <= 510, no wrap
    const orb_us2 lo = __builtin_elementwise_sub_sat(V, T);     // max(v - t, 0)
    const orb_us2 a = __builtin_bit_cast(orb_us2, p0), b = __builtin_bit_cast(orb_us2, p4);
    const orb_us2 c = __builtin_bit_cast(orb_us2, p8), d = __builtin_bit_cast(orb_us2, p12);
    // Some circularly adjacent pair (0,4), (4,8), (8,12), (12,0) is brighter
    // than hi <=> (0 or 8) and (4 or 12) are <=> min(max(p0, p8),
    // max(p4, p12)) > hi: opposite points never pair up.  Darker likewise
    // with max(min, min) < lo.  Nonzero lane <=> the pixel may pass.
    const orb_us2 M = __builtin_elementwise_min(__builtin_elementwise_max(a, c), __builtin_elementwise_max(b, d));
    const orb_us2 m = __builtin_elementwise_max(__builtin_elementwise_min(a, c), __builtin_elementwise_min(b, d));
    const orb_us2 r = __builtin_elementwise_max(__builtin_elementwise_sub_sat(M, hi), __builtin_elementwise_sub_sat(lo, m));
    return __builtin_bit_cast(unsigned, r);
}

// Bytes 0 and 2 (lo_bytes) or 1 and 3 (hi_bytes) of a dword, zero-extended
// into the two 16-bit lanes.
__device__ __forceinline__ unsigned lo_bytes(unsigned x) { return __builtin_amdgcn_perm(0u, x, 0x0C020C00u); }
__device__ __forceinline__ unsigned hi_bytes(unsigned x) { return __builtin_amdgcn_perm(0u, x, 0x0C030C01u); }

// Four adjacent pixels of cv::resize INTER_LINEAR 8U from the source rows at
// byte offsets o0, o1 of base (the LDS window, or the level in global memory
// with 32-bit offsets): horizontal 11-bit taps (a0 | a1 << 16) per column,
// vertical taps yb = b0 | b1 << 16, rounding shift by 22.
// - All terms are non-negative, so the shift is done unsigned.  (A signed
//   shift with a [0, 255] clamp of lanes 0 and 1 is lowered by the gfx950
//   backend to v_ashr_pk_u8_i32, whose upper result half then leaked into
//   bytes 2 and 3.)
// - Every factor is below 2^24 (taps <= 2048, S <= 255 * 4096), so the
//   products are 24-bit multiplies (v_mul_u32_u24, full rate) rather than
//   32-bit ones (v_mul_lo_u32, quarter rate).
typedef unsigned short orb_u16x2 __attribute__((ext_vector_type(2)));

// The vertical pass on the horizontal sums S0, S1 (each <= 255 * 2049) with
// the row taps b0, b1, rounded as OpenCV's generic scalar code does
// ((S0*b0 + S1*b1 + 2^21) >> 22) or as its SSE2 kernel VResizeLinearVec_32s8u
// does (S >> 4 as int16, _mm_mulhi_epi16 by the taps, + 2, >> 2).  Neither
// needs a clamp or a saturation: every tap pair sums to 2047..2049 (two
// roundings of values summing to 2048), so the scalar form is at most
// (255 * 2049 * 2049 + 2^21) >> 22 = 255, the SSE2 one at most
// (2 * ((32655 * 2049) >> 17) + 2) >> 2 = 255, and S >> 4 <= 32655 fits int16.
// MODE 0: scalar, 1: SSE2, 2: SSE2 where sse != 0.
template <int MODE>
__device__ __forceinline__ unsigned vresize(unsigned S0, unsigned S1, unsigned b0, unsigned b1, unsigned sse)
{
    const unsigned vc = (__umul24(S0, b0) + __umul24(S1, b1) + (1u << 21)) >> 22;
    const unsigned vs = ((__umul24(S0 >> 4, b0) >> 16) + (__umul24(S1 >> 4, b1) >> 16) + 2u) >> 2;
    return MODE == 0 ? vc : (MODE == 1 ? vs : (sse ? vs : vc));
}

// High 32 bits of a 24 x 24-bit product (v_mul_hi_u32_u24, full rate): the
// operands are masked to 24 bits, which is what lets the compiler pick it.
__device__ __forceinline__ unsigned mulhi_u24(unsigned a, unsigned b)
{
    return (unsigned)(((unsigned long long)(a & 0xFFFFFFu) * (unsigned long long)(b & 0xFFFFFFu)) >> 32);
}

template <int M>
struct ModeTag {
    static constexpr int value = M;
};

template <typename BytePtr>
__device__ __forceinline__ unsigned resize4(BytePtr base, unsigned o0, unsigned o1, const int* sxk, const int* nxk,
                                            const int* tap, int yb, const unsigned* sse)
{
    const unsigned b0 = (unsigned)yb & 0xFFFFu, b1 = (unsigned)yb >> 16;
    unsigned d = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const unsigned a0 = (unsigned)tap[k] & 0xFFFFu, a1 = (unsigned)tap[k] >> 16;
        const unsigned S0 = __umul24(base[o0 + sxk[k]], a0) + __umul24(base[o0 + nxk[k]], a1);
        const unsigned S1 = __umul24(base[o1 + sxk[k]], a0) + __umul24(base[o1 + nxk[k]], a1);
        d |= vresize<2>(S0, S1, b0, b1, sse[k]) << (8 * k);
    }
    return d;
}

// LDS that the phases use one after another: the resize source window
// (phase 1), the horizontal blur sums (phase 2), then the FAST queue and the
// tile's candidate list (phases 3-5).
#define LW_MAX 2656                  // source window dwords: scale <= ~1.9
#define LS_K 6                       // window dwords per thread read before the first wait
template <int TH>
union LevelScratch {
    unsigned win[LW_MAX];
    uint4 hs[TH + 6][LT_W / 4];
    struct {
        unsigned short queue[(TH + 2) * (LT_W + 2)];
        unsigned kloc[LK_CAP];       // tile-local candidate: ly | lx | score | retry
    } f;
};

template <int TH>
__global__ __launch_bounds__(256) void level_kernel(const ImageJob* __restrict__ jobs,
                                                    const RunTable* __restrict__ rt, int level)
{
    __shared__ __align__(16) unsigned pxw[(TH + 2 * LT_HALO)][LP_WD];   // level pixels, tile + 4-px halo
    __shared__ unsigned mmw[(TH + 2)][LP_WD];         // FAST strength, tile + 1-px ring
    __shared__ LevelScratch<TH> u;
    __shared__ int xs0[LP_W], xa[LP_W], ys0[(TH + 2 * LT_HALO)], yb[(TH + 2 * LT_HALO)];
    __shared__ int s_qn, s_nq, s_kn, s_base;
    __shared__ int s_col[LT_W], s_row[TH];      // FAST cell of each tile column / row, edge flags
    const uint8_t* px = (const uint8_t*)pxw;
    uint8_t* mm = (uint8_t*)mmw;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // Every wave finds its tile itself (uniform scalar loads), so the first
    // barrier comes only after the tile's memory reads have been issued; the
    // counters cleared here are first used after that barrier.
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int tpj = rt->tiles_per_job[level];
    const int job = tpj > 0 ? blk / tpj : find_segment(rt->tile_start[level], rt->njobs, blk);
    if (tid == 0) {
        s_qn = 0;
        s_nq = 0;
        s_kn = 0;
    }
    const ImageJob& J = jobs[job];
    const LevelJob& L = J.lv[level];
    const int t = blk - rt->tile_start[level][job];
    const int tx0 = (t % L.tiles_x) * LT_W;
    const int ty0 = (t / L.tiles_x) * TH;
    const int w = L.w, h = L.h, pitch = L.pitch;
    const bool interior = tx0 >= LT_HALO && tx0 + LT_W + LT_HALO <= w && ty0 >= LT_HALO &&
                          ty0 + TH + LT_HALO <= h;

    // FAST detection area of the level.
    const int ax0 = ORB_EDGE, ay0 = ORB_EDGE;
    const int ax1 = min(ORB_EDGE + L.nCols * L.wCell, L.maxBorderX - 3);
    const int ay1 = min(ORB_EDGE + L.nRows * L.hCell, L.maxBorderY - 3);
    // Per tile column: its FAST cell column cj, and whether the pixels left
    // and right of it fall outside that cell's window (bits 16, 17); rows
    // likewise.  Built while the tile's first reads are in flight; read
    // after later barriers.
    auto cell_tables = [&]() {
        if (tid < LT_W + TH) {
            const bool isc = tid < LT_W;
            const int z = isc ? tx0 + tid : ty0 + (tid - LT_W);
            const int cs = isc ? L.wCell : L.hCell, a0 = isc ? ax0 : ay0, a1 = isc ? ax1 : ay1;
            const int cz = (int)__umulhi((unsigned)max(z - ORB_EDGE, 0), isc ? L.wrcp : L.hrcp);   // / cs
            const int c0 = max(ORB_EDGE + cz * cs, a0), c1 = min(ORB_EDGE + (cz + 1) * cs, a1);
            const int info = cz | ((z - 1 < c0) << 16) | ((z + 1 >= c1) << 17);
            if (isc) s_col[tid] = info;
            else s_row[tid - LT_W] = info;
        }
    };

    // A resized dword of the tile proper (px row ey, dword c) also goes
    // straight to the level image.
    auto level_out = [&](int ey, int c, unsigned d) {
        const int y = ty0 + ey - LT_HALO, x0 = tx0 + 4 * (c - 1);
        if (ey >= LT_HALO && ey < LT_HALO + TH && c >= 1 && c <= LT_W / 4 && y < h) {
            uint8_t* dst = L.lvl + (unsigned)(y * pitch + x0);
            if (x0 + 3 < w) {
                gstore4(dst, d);
            } else {
                for (int k = 0; k < 4; k++)
                    if (x0 + k < w) dst[k] = (uint8_t)(d >> (8 * k));
            }
        }
    };

    // ---- 1. pixels with halo, one dword (4 columns) per slot
    if (level == 0) {
        const uint8_t* src = J.src;
        const int sp = J.src_pitch;
        if (interior) {
            // The 720 dwords, 3 per thread, read before any wait (indexes past
            // the end repeat the last dword); aligned when the input sits in
            // level 0, whose rows are 64-byte aligned.
            unsigned d[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int e = min(tid + 256 * k, (TH + 2 * LT_HALO) * LP_WD - 1);
                const int ey = e / LP_WD, ex = (e - ey * LP_WD) * 4;
                const unsigned off = (unsigned)((ty0 - LT_HALO + ey) * sp + (tx0 - LT_HALO + ex));
                d[k] = (sp & 3) == 0 ? gload4(src + off) : gload4u(src + off);
            }
            cell_tables();
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int e = tid + 256 * k;
                if (e < (TH + 2 * LT_HALO) * LP_WD) pxw[e / LP_WD][e % LP_WD] = d[k];
            }
        } else {
            cell_tables();
            for (int e = tid; e < (TH + 2 * LT_HALO) * LP_WD; e += 256) {
                const int ey = e / LP_WD, ex = (e - ey * LP_WD) * 4;
                const int y = reflect101(ty0 - LT_HALO + ey, h);
                const uint8_t* r = src + (size_t)y * sp;
                unsigned d = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) d |= (unsigned)r[reflect101(tx0 - LT_HALO + ex + k, w)] << (8 * k);
                pxw[ey][ex >> 2] = d;
            }
        }
    } else {
        // cv::resize INTER_LINEAR 8U fixed point from level l-1: the source
        // index and 11-bit taps of each column and row come from the level's
        // tables (computed on the host with OpenCV's double/float statements).
        const LevelJob& P = J.lv[level - 1];
        const int sw = P.w, sp = P.pitch;
        // The source rectangle the tile reads (the level's per-tile-column and
        // per-tile-row tables), in whole aligned dwords (level rows start
        // 64-byte aligned, so no read leaves its row's pitch).  It is staged
        // in LDS when it fits, in the same memory round trip as the taps;
        // otherwise (scale factors near 2 and above) the taps are read from
        // global memory directly.
        const uint2 xw = cload8(L.xwin + t % L.tiles_x), yw = cload8(L.ywin + t / L.tiles_x);
        const int wx0 = (int)xw.x & ~3, wy0 = (int)yw.x;
        const int nwd = ((int)xw.y >> 2) - (wx0 >> 2) + 1, nwr = (int)yw.y - wy0 + 1;
        const bool staged = nwd <= 64 && nwd * nwr <= LW_MAX;
        // All reads are issued before the first wait: one tap table entry per
        // thread (columns: threads 0..71, rows: 128..167) and the first
        // LS_K * 256 window dwords, in row-major order (row = e / nwd by a
        // 20-bit reciprocal, exact for e * nwd < 2^20).
        // (The row table follows the column table: ytab == xtab + w.)
        const int ti = tid < 128 ? reflect101(tx0 - LT_HALO + min(tid, LP_W - 1), w)
                                 : w + reflect101(ty0 - LT_HALO + min(tid - 128, (TH + 2 * LT_HALO) - 1), h);
        const uint2 tv = gload8((const uint8_t*)(L.xtab + ti));
        const int nst = staged ? nwd * nwr : 0;
        const unsigned rcp = ((1u << 20) + (unsigned)nwd - 1u) / (unsigned)nwd;
        const uint8_t* srow = P.lvl + (size_t)wy0 * sp + wx0;
        unsigned wv[LS_K];
        if (nst > 0) {
            // unconditional reads (indexes past the window repeat its last
            // dword), so that none of them waits for another
#pragma unroll
            for (int k = 0; k < LS_K; k++) {
                const int e = min(tid + 256 * k, nst - 1);
                const int r = (int)(__umul24((unsigned)e, rcp) >> 20);
                wv[k] = gload4(srow + (unsigned)(r * sp + 4 * (e - r * nwd)));
            }
        }
        cell_tables();
        if (tid < LP_W) {
            xs0[tid] = (int)tv.x;
            xa[tid] = (int)tv.y;
        } else if (tid >= 128 && tid < 128 + (TH + 2 * LT_HALO)) {
            ys0[tid - 128] = (int)tv.x;
            yb[tid - 128] = (int)tv.y;
        }
#pragma unroll
        for (int k = 0; k < LS_K; k++)
            if (tid + 256 * k < nst) u.win[tid + 256 * k] = wv[k];
        for (int e = tid + 256 * LS_K; e < nst; e += 256) {
            const int r = (int)(__umul24((unsigned)e, rcp) >> 20);
            u.win[e] = gload4(srow + (unsigned)(r * sp + 4 * (e - r * nwd)));
        }
        __syncthreads();
        // Each of 252 threads owns one 4-column group (taps kept in
        // registers) and rows r0, r0 + 14, r0 + 28 of the tile.
        if (tid < 14 * LP_WD) {
            const int c = tid % LP_WD, r0 = tid / LP_WD;
            const int xoff = staged ? wx0 : 0;
            int sxk[4], nxk[4], tap[4];
            unsigned sse[4];                       // SSE2 rounding for the column (xtab bit 30)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int e = xs0[4 * c + k];
                const int sx = e & 0x3FFFFFFF;
                sse[k] = (unsigned)e >> 30;
                sxk[k] = sx - xoff;
                nxk[k] = (sx + 1 < sw ? sx + 1 : sx) - xoff;
                tap[k] = xa[4 * c + k];
            }
            // Rounding of the tile's columns (reflected halo included): all
            // scalar, all SSE2, or per column where the tile reaches the last
            // columns the SSE2 kernel leaves to the scalar pass.
            const int maxcol = max(min(tx0 + LT_W + LT_HALO - 1, w - 1), LT_HALO);
            const int rmode = L.rs_end == 0 ? 0 : (maxcol < L.rs_end ? 1 : 2);
            // The four pixels' source bytes usually lie within 8 bytes (scale
            // factors up to ~2.3; reflected halo columns may run backwards, so
            // the span is taken over all of them): then a source row is three
            // aligned dword reads realigned by two v_alignbyte, each pixel's
            // tap pair (p[sx], p[nx]) one v_perm into 16-bit lanes, and S one
            // v_dot2_u32_u16 with the packed taps (a0 | a1 << 16).
            int b0 = sxk[0], b1 = nxk[0];
#pragma unroll
            for (int k = 1; k < 4; k++) {
                b0 = min(b0, sxk[k]);
                b1 = max(b1, nxk[k]);
            }
            if (staged && b1 - b0 <= 7) {
                const unsigned* wd = u.win + (b0 >> 2);
                const unsigned shb = (unsigned)(b0 & 3);
                const int ws = nwd;                        // window row stride in dwords
                unsigned sel[4];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    sel[k] = 0x0C000C00u | ((unsigned)(nxk[k] - b0) << 16) | (unsigned)(sxk[k] - b0);
                auto rows = [&](auto mode) {
                    constexpr int MODE = decltype(mode)::value;
                    for (int ey = r0; ey < (TH + 2 * LT_HALO); ey += 14) {
                        const unsigned sy01 = (unsigned)ys0[ey];
                        const unsigned* q0 = wd + __umul24((sy01 & 0xFFFFu) - wy0, ws);
                        const unsigned* q1 = wd + __umul24((sy01 >> 16) - wy0, ws);
                        const unsigned w00 = __builtin_amdgcn_alignbyte(q0[1], q0[0], shb);
                        const unsigned w01 = __builtin_amdgcn_alignbyte(q0[2], q0[1], shb);
                        const unsigned w10 = __builtin_amdgcn_alignbyte(q1[1], q1[0], shb);
                        const unsigned w11 = __builtin_amdgcn_alignbyte(q1[2], q1[1], shb);
                        const unsigned ybv = (unsigned)yb[ey];
                        unsigned d = 0;
                        if constexpr (MODE == 1) {
                            // SSE2 rounding with the horizontal taps scaled by
                            // 16: S16 = 16 S, so S16 & 0xFFFF00 = (S >> 4) << 8
                            // (S16 < 2^23), and the high word of its product
                            // with b << 8 is ((S >> 4) * b) >> 16 in one
                            // v_mul_hi_u32_u24.  t = h0 + h1 + 2 <= 1023 fits
                            // a 16-bit lane: the >> 2 is done two pixels at a
                            // time and the four bytes gathered by one v_perm.
                            const unsigned bs0 = (ybv << 8) & 0xFFFF00u, bs1 = (ybv >> 8) & 0xFFFF00u;
                            unsigned t[4];
#pragma unroll
                            for (int k = 0; k < 4; k++) {
                                const orb_u16x2 tk = __builtin_bit_cast(orb_u16x2, (unsigned)tap[k] << 4);
                                const unsigned S0 = __builtin_amdgcn_udot2(
                                    __builtin_bit_cast(orb_u16x2, __builtin_amdgcn_perm(w01, w00, sel[k])), tk, 0u, false);
                                const unsigned S1 = __builtin_amdgcn_udot2(
                                    __builtin_bit_cast(orb_u16x2, __builtin_amdgcn_perm(w11, w10, sel[k])), tk, 0u, false);
                                t[k] = mulhi_u24(S0 & 0xFFFF00u, bs0) + mulhi_u24(S1 & 0xFFFF00u, bs1) + 2u;
                            }
                            const orb_u16x2 two = {2, 2};
                            const unsigned u01 = __builtin_bit_cast(unsigned, __builtin_bit_cast(orb_u16x2, t[0] | (t[1] << 16)) >> two);
                            const unsigned u23 = __builtin_bit_cast(unsigned, __builtin_bit_cast(orb_u16x2, t[2] | (t[3] << 16)) >> two);
                            d = __builtin_amdgcn_perm(u23, u01, 0x06040200u);
                        } else {
                            const unsigned bv0 = ybv & 0xFFFFu, bv1 = ybv >> 16;
#pragma unroll
                            for (int k = 0; k < 4; k++) {
                                const orb_u16x2 tk = __builtin_bit_cast(orb_u16x2, (unsigned)tap[k]);
                                const unsigned S0 = __builtin_amdgcn_udot2(
                                    __builtin_bit_cast(orb_u16x2, __builtin_amdgcn_perm(w01, w00, sel[k])), tk, 0u, false);
                                const unsigned S1 = __builtin_amdgcn_udot2(
                                    __builtin_bit_cast(orb_u16x2, __builtin_amdgcn_perm(w11, w10, sel[k])), tk, 0u, false);
                                d |= vresize<MODE>(S0, S1, bv0, bv1, sse[k]) << (8 * k);
                            }
                        }
                        pxw[ey][c] = d;
                        level_out(ey, c, d);
                    }
                };
                if (rmode == 1) rows(ModeTag<1>{});
                else if (rmode == 0) rows(ModeTag<0>{});
                else rows(ModeTag<2>{});
            } else if (staged) {
                const uint8_t* wb = (const uint8_t*)u.win;
                const int ws = nwd * 4;
                for (int ey = r0; ey < (TH + 2 * LT_HALO); ey += 14) {
                    const unsigned sy01 = (unsigned)ys0[ey];
                    const unsigned d = resize4(wb, __umul24((sy01 & 0xFFFFu) - wy0, ws),
                                               __umul24((sy01 >> 16) - wy0, ws), sxk, nxk, tap, yb[ey], sse);
                    pxw[ey][c] = d;
                    level_out(ey, c, d);
                }
            } else {
                for (int ey = r0; ey < (TH + 2 * LT_HALO); ey += 14) {
                    const unsigned sy01 = (unsigned)ys0[ey];
                    const unsigned d = resize4((const __attribute__((address_space(1))) uint8_t*)P.lvl,
                                               __umul24(sy01 & 0xFFFFu, sp), __umul24(sy01 >> 16, sp), sxk, nxk, tap,
                                               yb[ey], sse);
                    pxw[ey][c] = d;
                    level_out(ey, c, d);
                }
            }
        }
    }
    __syncthreads();

    // ---- 2. the level (copied out) and its 7x7 blur.  Horizontal taps as
    // two v_dot4_u32_u8 per output on byte quads (v_alignbyte); the vertical
    // pass on the row sums, symmetric taps.
    const int gc = tid & 15;                      // 4-column group of the tile
    // Level 0 copied out when the input is not level 0 itself (levels >= 1
    // are written by the resize).
    if (level == 0 && J.src != L.lvl) {
        for (int r = tid >> 4; r < TH; r += 16) {
            const int y = ty0 + r, x0 = tx0 + 4 * gc;
            if (y >= h) continue;
            const unsigned d = pxw[r + LT_HALO][gc + 1];
            uint8_t* dst = L.lvl + (unsigned)(y * pitch + x0);
            if (x0 + 3 < w) {
                gstore4(dst, d);
            } else {
                for (int k = 0; k < 4; k++)
                    if (x0 + k < w) dst[k] = (uint8_t)(d >> (8 * k));
            }
        }
    }
    const unsigned T1 = (unsigned)c_gauss[0] | (c_gauss[1] << 8) | (c_gauss[2] << 16) | ((unsigned)c_gauss[3] << 24);
    const unsigned T2 = (unsigned)c_gauss[4] | (c_gauss[5] << 8) | (c_gauss[6] << 16);
    // Two adjacent 4-column groups per task (they share two source dwords).
    for (int e = tid; e < (TH + 6) * (LT_W / 8); e += 256) {
        const int rr = e >> 3, c = 2 * (e & 7);  // blur row rr <-> px row rr+1
        const unsigned* pr = &pxw[rr + 1][c];
        const unsigned D[4] = {pr[0], pr[1], pr[2], pr[3]};
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const unsigned D0 = D[g], D1 = D[g + 1], D2 = D[g + 2];
            uint4 o;
            o.x = __builtin_amdgcn_udot4(__builtin_amdgcn_alignbyte(D1, D0, 1), T1,
                                         __builtin_amdgcn_udot4(__builtin_amdgcn_alignbyte(D2, D1, 1), T2, 0u, false),
                                         false);
            o.y = __builtin_amdgcn_udot4(__builtin_amdgcn_alignbyte(D1, D0, 2), T1,
                                         __builtin_amdgcn_udot4(__builtin_amdgcn_alignbyte(D2, D1, 2), T2, 0u, false),
                                         false);
            o.z = __builtin_amdgcn_udot4(__builtin_amdgcn_alignbyte(D1, D0, 3), T1,
                                         __builtin_amdgcn_udot4(__builtin_amdgcn_alignbyte(D2, D1, 3), T2, 0u, false),
                                         false);
            o.w = __builtin_amdgcn_udot4(D1, T1, __builtin_amdgcn_udot4(D2, T2, 0u, false), false);
            u.hs[rr][c + g] = o;
        }
    }
    __syncthreads();
    {
        // Row sums are <= 255 * 257 and taps < 256: 24-bit multiplies.
        const unsigned k0 = (unsigned)c_gauss[3], k1 = (unsigned)c_gauss[4], k2 = (unsigned)c_gauss[5],
                       k3 = (unsigned)c_gauss[6];
        // rows r .. r + RP - 1 of the tile from 6 + RP row-sum vectors
        constexpr int RP = TH / 16;
        static_assert(RP * 16 == TH && 16 * (LT_W / 4) == 256, "RP rows and one column group per thread");
        const int r = (tid >> 4) * RP, x0 = tx0 + 4 * gc;
        // Column rounding: columns < bs_end as OpenCV's SSE2 column kernel
        // (SymmColumnVec_32s8u: the taps as floats k / 2^16, the sum rounded
        // half to even, which for these exact sums is (acc + 0x7FFF + bit 16
        // of acc) >> 16), the others as its scalar one ((acc + 2^15) >> 16).
        unsigned rb[4];
#pragma unroll
        for (int k = 0; k < 4; k++) rb[k] = x0 + k < L.bs_end ? 1u : 0u;
        uint4 q[6 + RP];
#pragma unroll
        for (int i = 0; i < 6 + RP; i++) q[i] = u.hs[r + i][gc];
#pragma unroll
        for (int hh = 0; hh < RP; hh++) {
            const int y = ty0 + r + hh;
            if (y >= h) continue;
            unsigned out = 0;
#define ORB_VBLUR(comp, k)                                                                              \
            {                                                                                            \
                const unsigned acc = __umul24(q[hh + 3].comp, k0) +                                      \
                                     __umul24(q[hh + 2].comp + q[hh + 4].comp, k1) +                     \
                                     __umul24(q[hh + 1].comp + q[hh + 5].comp, k2) +                     \
                                     __umul24(q[hh].comp + q[hh + 6].comp, k3);                          \
                const unsigned v = (acc + (1u << 15) - rb[k] + ((acc >> 16) & rb[k])) >> 16;            \
                out |= (v > 255u ? 255u : v) << (8 * (k));                                              \
            }
            ORB_VBLUR(x, 0) ORB_VBLUR(y, 1) ORB_VBLUR(z, 2) ORB_VBLUR(w, 3)
#undef ORB_VBLUR
            uint8_t* db = L.blur + (unsigned)(y * pitch + x0);
            if (x0 + 3 < w) {
                gstore4(db, out);
            } else {
                for (int k = 0; k < 4; k++)
                    if (x0 + k < w) db[k] = (uint8_t)(out >> (8 * k));
            }
        }
    }
    __syncthreads();                              // hs is reused below

    // ---- 3. FAST strength on the tile plus a 1-px ring, detection areas
    // only.  The cardinal pre-test runs on 12-pixel groups (two pixels per
    // 16-bit lane pair); survivors are compacted, then scored exactly.
    const int tA = min(max(J.iniTh, 0), 255), tB = min(max(J.minTh, 0), 255);
    const int tmin = min(tA, tB);
    const unsigned t2 = (unsigned)tmin | ((unsigned)tmin << 16);
    // tile plus its 1-px ring wholly inside the detection area
    const bool tile_in = tx0 - 1 >= ax0 && tx0 + LT_W < ax1 && ty0 - 1 >= ay0 && ty0 + TH < ay1;
    // Lanes: rows tile -1..32 and 12-column groups g = 0..5 (px dwords 3g ..
    // 3g + 2, tile columns 12g - 4 .. 12g + 7): 204 lanes, one pass.  Group 0
    // only contributes its pixels 3.. (tile columns -1..) and group 5 its
    // pixels ..8 (columns ..64): the 1-px ring.  Their missing neighbour
    // dwords (columns -8 and 68 onwards) feed only the bytes masked off.
    static_assert((TH + 2) * 6 <= 256, "one pre-test pass");
    static_assert(LP_WD == 18, "6 lanes of 3 dwords cover a strength-map row");
    {
        unsigned cand = 0u;                       // bit k: pixel k of the group
        int my = 0, g = 0;
        if (tid < (TH + 2) * 6) {
            my = tid / 6;
            g = tid - my * 6;
            const int y = ty0 - 1 + my;
            const int py = my + 3;                // px row of tile row y
            const int x = tx0 + 12 * g - 4;
            const unsigned* vr = &pxw[py][3 * g];
            const unsigned* ur = &pxw[py - 3][3 * g];
            const unsigned* dr = &pxw[py + 3][3 * g];
            // The lanes tile the strength map exactly (34 rows x 18 dwords):
            // each clears its 3 dwords; candidates get their strength below.
            mmw[my][3 * g] = 0u;
            mmw[my][3 * g + 1] = 0u;
            mmw[my][3 * g + 2] = 0u;
            const unsigned Vx = vr[0], Vy = vr[1], Vz = vr[2];
            const unsigned Lw = g > 0 ? vr[-1] : 0u;
            const unsigned Rw = g < 5 ? vr[3] : 0u;
            // Even / odd pixels of each dword in 16-bit lanes (lo / hi =
            // pixels 0, 2 / 1, 3 of the dword); the neighbours 3 columns
            // right (p4) and left (p12) are byte selections of the row.
            const unsigned Vxl = lo_bytes(Vx), Vxh = hi_bytes(Vx), Vyl = lo_bytes(Vy), Vyh = hi_bytes(Vy);
            const unsigned Vzl = lo_bytes(Vz), Vzh = hi_bytes(Vz);
            const unsigned m02a = fast_maybe2(Vxl, lo_bytes(dr[0]), __builtin_amdgcn_perm(Vy, Vx, 0x0C050C03u),
                                              lo_bytes(ur[0]), hi_bytes(Lw), t2);        // columns 0, 2
            const unsigned m13a = fast_maybe2(Vxh, hi_bytes(dr[0]), Vyl, hi_bytes(ur[0]),
                                              __builtin_amdgcn_perm(Vx, Lw, 0x0C040C02u), t2);   // 1, 3
            const unsigned m02b = fast_maybe2(Vyl, lo_bytes(dr[1]), __builtin_amdgcn_perm(Vz, Vy, 0x0C050C03u),
                                              lo_bytes(ur[1]), Vxh, t2);                 // 4, 6
            const unsigned m13b = fast_maybe2(Vyh, hi_bytes(dr[1]), Vzl, hi_bytes(ur[1]),
                                              __builtin_amdgcn_perm(Vy, Vx, 0x0C040C02u), t2);   // 5, 7
            const unsigned m02c = fast_maybe2(Vzl, lo_bytes(dr[2]), __builtin_amdgcn_perm(Rw, Vz, 0x0C050C03u),
                                              lo_bytes(ur[2]), Vyh, t2);                 // 8, 10
            const unsigned m13c = fast_maybe2(Vzh, hi_bytes(dr[2]), lo_bytes(Rw), hi_bytes(ur[2]),
                                              __builtin_amdgcn_perm(Vz, Vy, 0x0C040C02u), t2);   // 9, 11
            // pixels k in [klo, khi): detection area, and the ring rule
            unsigned mask;
            if (tile_in) {
                mask = g == 0 ? 0xFF8u : (g == 5 ? 0x1FFu : 0xFFFu);
            } else {
                const int klo = min(max(max(ax0 - x, g == 0 ? 3 : 0), 0), 12);
                const int khi = max(min(min(ax1 - x, g == 5 ? 9 : 12), 12), 0);
                const bool yin = y >= ay0 && y < ay1;
                mask = (yin && klo < khi) ? ((1u << khi) - (1u << klo)) : 0u;
            }
            // each 16-bit lane clamped to 0 / 1: pixel k's bit at k (k = 0, 1,
            // 4, 5, 8, 9) or k + 14 (k = 2, 3, 6, 7, 10, 11)
            // (1 | 1 << 16 hidden from the optimiser, which otherwise turns
            // the clamp into per-lane compares and selects)
            unsigned one2 = 0x00010001u;
            asm volatile("" : "+s"(one2));
            const orb_us2 one = __builtin_bit_cast(orb_us2, one2);
#define ORB_BIT01(m) __builtin_bit_cast(unsigned, __builtin_elementwise_min(__builtin_bit_cast(orb_us2, m), one))
            const unsigned x01 = ORB_BIT01(m02a) | (ORB_BIT01(m13a) << 1) | (ORB_BIT01(m02b) << 4) |
                                 (ORB_BIT01(m13b) << 5) | (ORB_BIT01(m02c) << 8) | (ORB_BIT01(m13c) << 9);
#undef ORB_BIT01
            cand = (x01 | (x01 >> 14)) & mask;
        }
        // Append: the wave's exclusive prefix of the per-lane counts (0..12)
        // from four ballots of the count's bits, one LDS atomic per wave.
        const int cnt = __popc(cand);
        const unsigned long long b0 = __ballot(cnt & 1), b1 = __ballot(cnt & 2), b2 = __ballot(cnt & 4),
                                 b3 = __ballot(cnt & 8);
        const int tot = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2) + 8 * __popcll(b3);
        if (tot) {
            int base = 0;
            if (lane == 0) base = atomicAdd(&s_qn, tot);
            base = __builtin_amdgcn_readfirstlane(base);     // every lane is active here
            int pos = base + lanes_below(b0) + 2 * lanes_below(b1) + 4 * lanes_below(b2) + 8 * lanes_below(b3);
            // queue entry: m-row << 8 | byte column in mmw (tile col + 4)
            const unsigned ebase = ((unsigned)my << 8) | (unsigned)(12 * g);
            for (unsigned m = cand; m; m &= m - 1)
                u.f.queue[pos++] = (unsigned short)(ebase + (unsigned)__builtin_ctz(m));
        }
    }
    __syncthreads();
    // Exact strengths of the queued pixels, two per thread before one barrier.
    // Pixels of the tile proper that beat the lower threshold are compacted
    // again, in place (a slot is only rewritten after the barrier that ends
    // the reads of its chunk of 512), so that suppression runs on dense lanes.
    const int qn = s_qn;
    for (int q0 = 0; q0 < qn; q0 += 512) {
        int e[2] = {0, 0};
        bool keep[2] = {false, false};
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int q = q0 + 256 * h + tid;
            if (q < qn) {
                e[h] = u.f.queue[q];
                const int my = e[h] >> 8, col = e[h] & 255;
                const int sv = fast_strength(px + (my + 3) * LP_W + col, LP_W);
                mm[my * LP_W + col] = (uint8_t)sv;
                keep[h] = sv > tmin && my >= 1 && my <= TH && col >= 4 && col < 4 + LT_W;
            }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const unsigned long long bal = __ballot(keep[h]);
            if (bal) {
                int base = 0;
                if (lane == 0) base = atomicAdd(&s_nq, __popcll(bal));
                base = __builtin_amdgcn_readfirstlane(base);     // every lane is active here
                if (keep[h]) u.f.queue[base + lanes_below(bal)] = (unsigned short)e[h];
            }
        }
    }
    __syncthreads();

    // ---- 4. per-cell non-max suppression at both thresholds.  A neighbour's
    // score at threshold t is (m > t ? m - 1 : 0) for its strength m, 0 outside
    // the cell's window; it grows with m, so the largest neighbour score is
    // the score of the largest in-window neighbour strength M.  The pixel is
    // kept at t iff mc > t and mc - 1 > score(M) (which is >= 0, so this also
    // beats the out-of-window zeros).
    const int nq = s_nq;
    for (int q0 = 0; q0 < nq; q0 += 256) {
        const int q = q0 + tid;
        bool keepA = false, keepB = false;
        unsigned loc = 0u;
        if (q < nq) {
            const int e = u.f.queue[q];
            const int ly = (e >> 8) - 1, lx = (e & 255) - 4;
            const uint8_t* mrow = mm + (ly + 1) * LP_W + (lx + 4);
            const int mc = mrow[0];
            const int colv = s_col[lx], rowv = s_row[ly];
            int M = 0;
#pragma unroll
            for (int n = 0; n < 9; n++) {
                if (n == 4) continue;
                const int dx = n % 3 - 1, dy = n / 3 - 1;
                const bool out = (dx < 0 && (colv & 0x10000)) || (dx > 0 && (colv & 0x20000)) ||
                                 (dy < 0 && (rowv & 0x10000)) || (dy > 0 && (rowv & 0x20000));
                const int mq = mrow[dy * LP_W + dx];
                M = max(M, out ? 0 : mq);
            }
            keepA = mc > tA && mc - 1 > (M > tA ? M - 1 : 0);
            keepB = mc > tB && mc - 1 > (M > tB ? M - 1 : 0);
            loc = ((unsigned)ly << 6) | (unsigned)lx | ((unsigned)mc << 11);
            if (keepA) gstore1(L.cellflag + (rowv & 0xFFFF) * L.nCols + (colv & 0xFFFF), 1);
        }
        // Append (A entry, then B entry, per lane): the wave's exclusive
        // prefix from two ballots, one LDS atomic per wave.
        const unsigned long long ba = __ballot(keepA), bb = __ballot(keepB);
        const int tot = __popcll(ba) + __popcll(bb);
        if (tot) {
            int base = 0;
            if (lane == 0) base = atomicAdd(&s_kn, tot);
            base = __builtin_amdgcn_readfirstlane(base);     // every lane is active here
            int pos = base + lanes_below(ba) + lanes_below(bb);
            if (keepA) {
                if (pos < LK_CAP) u.f.kloc[pos] = loc;
                pos++;
            }
            if (keepB && pos < LK_CAP) u.f.kloc[pos] = loc | (1u << 19);
        }
    }
    __syncthreads();

    // ---- 5. append the tile's candidates to the level list as
    // (score << 32 | cell row, cell col, y, x in cell) [| retry]
    const int kn = s_kn;
    if (kn == 0) return;
    if (tid == 0) {
        s_base = atomicAdd(&J.counters[ORB_CNT_KEYS + level], kn);
        if (kn > LK_CAP || s_base + kn > L.key_cap) atomicOr(&J.counters[ORB_CNT_ERR], ORB_ERR_KEYS);
    }
    __syncthreads();
    const int base = s_base;
    for (int i = tid; i < min(kn, LK_CAP); i += 256) {
        if (base + i >= L.key_cap) break;
        const unsigned loc = u.f.kloc[i];
        const int x = tx0 + (int)(loc & 63u), y = ty0 + (int)((loc >> 6) & 31u);
        const int mc = (int)((loc >> 11) & 255u);
        const int cj = s_col[loc & 63u] & 0xFFFF, ci = s_row[(loc >> 6) & 31u] & 0xFFFF;
        const unsigned xl = (unsigned)(x - (ORB_MIN_BORDER + cj * L.wCell));
        const unsigned yl = (unsigned)(y - (ORB_MIN_BORDER + ci * L.hCell));
        const unsigned ok = ((unsigned)ci << 20) | ((unsigned)cj << 12) | (yl << 6) | xl;
        unsigned long long key = ((unsigned long long)(unsigned)(mc - 1) << 32) | ok;
        if (loc & (1u << 19)) key |= ORB_KEY_RETRY;
        gstore8(L.keys + base + i, key);
    }
}

// ================================================================ octree
// DistributeOctTree (ORBextractor.cc:688-1033) on one workgroup per level.
//
// The reference keeps a std::list of nodes.  Phase 1 repeats full passes,
// splitting every node with more than one key and pushing its non-empty
// children to the list front (n1..n4, so they read n4,n3,n2,n1); phase 2
// sorts the expandable nodes by (size, address) and splits from the largest
// until the quota is reached.  Here the list is an array in LDS that is
// rebuilt each pass with prefix sums; each candidate carries the index of the
// list node that holds it.  Node order, membership and retained keys are
// exactly the list's.  Ties in the phase-2 sort are broken by creation order
// (newer = larger), the oracle's fixed stand-in for the heap address.
#define OCT_T 1024
#define OCT_REG 4                    // candidates per thread held in registers

struct OctShared {
    int rx[2][ORB_LIST_CAP];     // x0 | x1 << 16
    int ry[2][ORB_LIST_CAP];     // y0 | y1 << 16
    int cnt[2][ORB_LIST_CAP];
    int seq[2][ORB_LIST_CAP];
    int cc[ORB_LIST_CAP][4];
    short map[ORB_LIST_CAP][4];
    short keep[ORB_LIST_CAP];
    unsigned long long sortk[ORB_LIST_CAP];
    unsigned long long best[ORB_LIST_CAP];
    int scan[OCT_T / 64 + 1];
    int vars[8];
};

__device__ __forceinline__ int block_scan(int v, int* sh, int* total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (wid == 0) {
        const int wsum = lane < OCT_T / 64 ? sh[lane] : 0;
        int s = wsum;
#pragma unroll
        for (int o = 1; o < OCT_T / 64; o <<= 1) {
            const int y = __shfl_up(s, o, 64);
            if (lane >= o) s += y;
        }
        if (lane < OCT_T / 64) sh[lane] = s - wsum;
        if (lane == OCT_T / 64 - 1) sh[OCT_T / 64] = s;
    }
    __syncthreads();
    const int r = x - v + sh[wid];
    *total = sh[OCT_T / 64];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void key_xy(unsigned long long key, int wCell, int hCell, int* x, int* y)
{
    const unsigned ok = (unsigned)key;
    const int xl = ok & 63, yl = (ok >> 6) & 63, j = (ok >> 12) & 255, i = (ok >> 20) & 255;
    *x = j * wCell + xl;
    *y = i * hCell + yl;
}

__device__ __forceinline__ int quad_of(int rx, int ry, int x, int y)
{
    const int x0 = rx & 0xFFFF, x1 = rx >> 16, y0 = ry & 0xFFFF, y1 = ry >> 16;
    const int mx = x0 + ((x1 - x0 + 1) >> 1), my = y0 + ((y1 - y0 + 1) >> 1);
    return (x < mx ? 0 : 1) + (y < my ? 0 : 2);
}

__device__ __forceinline__ void child_rect(int rx, int ry, int q, int* crx, int* cry)
{
    const int x0 = rx & 0xFFFF, x1 = rx >> 16, y0 = ry & 0xFFFF, y1 = ry >> 16;
    const int mx = x0 + ((x1 - x0 + 1) >> 1), my = y0 + ((y1 - y0 + 1) >> 1);
    const int cx0 = (q & 1) ? mx : x0, cx1 = (q & 1) ? x1 : mx;
    const int cy0 = (q & 2) ? my : y0, cy1 = (q & 2) ? y1 : my;
    *crx = cx0 | (cx1 << 16);
    *cry = cy0 | (cy1 << 16);
}

__global__ __launch_bounds__(OCT_T, 8) void octree_kernel(const ImageJob* __restrict__ jobs, int max_levels)
{
    __shared__ OctShared S;
    const int tid = threadIdx.x;
    const int jb = blockIdx.x / max_levels, level = blockIdx.x % max_levels;
    const ImageJob& J = jobs[jb];
    if (level >= J.nlevels) return;
    const LevelJob& L = J.lv[level];
    int* ctr = J.counters;
    const int K0 = min(ctr[ORB_CNT_KEYS + level], L.key_cap);
    const int N = L.quota;
    const int wC = L.wCell, hC = L.hCell;
    int* node = L.node;
    int total;

    // ---- per-cell retry (ORBextractor.cc:1111-1124): a cell keeps its
    // iniThFAST candidates when it has any, else its minThFAST candidates.
    // OCT_REG candidates per thread per round: their keys, then their cell
    // flags, are read together, and one scan places the survivors.  (The
    // candidates' order does not matter: the level kernels append them in
    // no fixed order, and every later step depends only on the set.)
    int K = 0;
    for (int c0 = 0; c0 < K0; c0 += OCT_T * OCT_REG) {
        unsigned long long key[OCT_REG];
        unsigned char flag[OCT_REG];
#pragma unroll
        for (int m = 0; m < OCT_REG; m++) key[m] = gload8u64(L.keys + min(c0 + tid + OCT_T * m, K0 - 1));
#pragma unroll
        for (int m = 0; m < OCT_REG; m++) {
            const unsigned ok = (unsigned)key[m];
            flag[m] = gload1(L.cellflag + ((ok >> 20) & 255) * L.nCols + ((ok >> 12) & 255));
        }
        int cnt = 0;
        bool v[OCT_REG];
#pragma unroll
        for (int m = 0; m < OCT_REG; m++) {
            v[m] = c0 + tid + OCT_T * m < K0 && ((key[m] & ORB_KEY_RETRY) ? flag[m] == 0 : flag[m] != 0);
            cnt += v[m] ? 1 : 0;
        }
        int pos = block_scan(cnt, S.scan, &total);
#pragma unroll
        for (int m = 0; m < OCT_REG; m++)
            if (v[m]) gstore8(L.keys2 + K + pos++, key[m] & ~ORB_KEY_RETRY);
        K += total;
    }
    __syncthreads();
    const unsigned long long* keys = L.keys2;

    // Each thread keeps the position and list node of its candidates
    // k = tid + OCT_T*m, m < OCT_REG, in registers for every pass; candidates
    // beyond OCT_REG * OCT_T (rare) keep their node in global memory.
    int kx[OCT_REG], ky[OCT_REG], kn[OCT_REG];
    unsigned long long rk[OCT_REG];
#pragma unroll
    for (int m = 0; m < OCT_REG; m++) rk[m] = K > 0 ? gload8u64(keys + min(tid + OCT_T * m, K - 1)) : 0ull;
#pragma unroll
    for (int m = 0; m < OCT_REG; m++) {
        kx[m] = ky[m] = kn[m] = 0;
        if (tid + OCT_T * m < K) key_xy(rk[m], wC, hC, &kx[m], &ky[m]);
    }
    // f(x, y, node&) on every candidate of this thread.
    auto for_keys = [&](auto&& f) {
#pragma unroll
        for (int m = 0; m < OCT_REG; m++)
            if (tid + OCT_T * m < K) f(kx[m], ky[m], kn[m]);
        for (int k = tid + OCT_T * OCT_REG; k < K; k += OCT_T) {
            int x, y;
            key_xy(keys[k], wC, hC, &x, &y);
            int nd = node[k];
            f(x, y, nd);
            node[k] = nd;
        }
    };

    // ---- roots (ORBextractor.cc:695-763)
    if (tid < ORB_MAX_ROOTS) S.cc[tid][0] = 0;
    __syncthreads();
    for_keys([&](int x, int, int& nd) {
        int r = (int)__fdiv_rn((float)x, L.hX);
        r = min(r, L.nIni - 1);
        nd = r;
        atomicAdd(&S.cc[r][0], 1);
    });
    __syncthreads();
    {
        const int c = tid < L.nIni ? S.cc[tid][0] : 0;
        const int pos = block_scan(c > 0 ? 1 : 0, S.scan, &total);
        if (tid < L.nIni && c > 0) {
            S.rx[0][pos] = L.root_x[tid] | (L.root_x[tid + 1] << 16);
            S.ry[0][pos] = 0 | ((L.maxBorderY - ORB_MIN_BORDER) << 16);
            S.cnt[0][pos] = c;
            S.seq[0][pos] = tid;
        }
        if (tid < L.nIni) S.keep[tid] = (short)pos;
    }
    __syncthreads();
    for_keys([&](int, int, int& nd) { nd = S.keep[nd]; });
    int cur = 0;
    int size = total;
    bool phase2 = false;
    __syncthreads();

    // ---- phase 1: full passes (ORBextractor.cc:779-893).  Every pass
    // that does not end the loop grows the list, which holds at most
    // ORB_LIST_CAP nodes, so a pass count beyond that is an internal error.
    for (int pass = 0;; pass++) {
        if (pass > ORB_LIST_CAP) {
            if (tid == 0) atomicOr(&ctr[ORB_CNT_ERR], ORB_ERR_LIST);
            size = 0;
            break;
        }
        const int prev = size;
        for (int i = tid; i < size; i += OCT_T) S.cc[i][0] = S.cc[i][1] = S.cc[i][2] = S.cc[i][3] = 0;
        __syncthreads();
        for_keys([&](int x, int y, int& nd) {
            const int i = nd;
            if (S.cnt[cur][i] >= 2) atomicAdd(&S.cc[i][quad_of(S.rx[cur][i], S.ry[cur][i], x, y)], 1);
        });
        __syncthreads();
        const int i = tid;
        const bool split = i < size && S.cnt[cur][i] >= 2;
        int nc = 0;
        if (split) nc = (S.cc[i][0] > 0) + (S.cc[i][1] > 0) + (S.cc[i][2] > 0) + (S.cc[i][3] > 0);
        const int kp = (i < size && !split) ? 1 : 0;
        int C, nkeep;
        const int ex_nc = block_scan(nc, S.scan, &C);
        const int ex_kp = block_scan(kp, S.scan, &nkeep);
        const int nsize = C + nkeep;
        int nexp = 0;
        if (nsize <= ORB_LIST_CAP) {
            const int nb = cur ^ 1;
            if (split) {
                const int base = C - ex_nc - nc;
                int r = 0;
                for (int q = 3; q >= 0; q--) {
                    const int c = S.cc[i][q];
                    if (c > 0) {
                        const int pos = base + r++;
                        int crx, cry;
                        child_rect(S.rx[cur][i], S.ry[cur][i], q, &crx, &cry);
                        S.rx[nb][pos] = crx;
                        S.ry[nb][pos] = cry;
                        S.cnt[nb][pos] = c;
                        S.seq[nb][pos] = ex_nc + (nc - r);   // creation order n1..n4
                        S.map[i][q] = (short)pos;
                        nexp += c >= 2;
                    } else {
                        S.map[i][q] = -1;
                    }
                }
            } else if (kp) {
                const int pos = C + ex_kp;
                S.rx[nb][pos] = S.rx[cur][i];
                S.ry[nb][pos] = S.ry[cur][i];
                S.cnt[nb][pos] = S.cnt[cur][i];
                S.seq[nb][pos] = S.seq[cur][i];
                S.keep[i] = (short)pos;
            }
        }
        int nToExpand;
        block_scan(nexp, S.scan, &nToExpand);
        if (nsize > ORB_LIST_CAP) {
            if (tid == 0) atomicOr(&ctr[ORB_CNT_ERR], ORB_ERR_LIST);
            size = 0;
            break;
        }
        for_keys([&](int x, int y, int& nd) {
            const int n0 = nd;
            nd = S.cnt[cur][n0] >= 2 ? S.map[n0][quad_of(S.rx[cur][n0], S.ry[cur][n0], x, y)] : S.keep[n0];
        });
        cur ^= 1;
        size = nsize;
        __syncthreads();
        if (size >= N || size == prev) break;
        if (size + nToExpand * 3 > N) {
            phase2 = true;
            break;
        }
    }

    // ---- phase 2: split the largest nodes until the quota (:908-991); as
    // in phase 1, every round that does not end the loop grows the list.
    for (int round = 0; phase2; round++) {
        if (round > ORB_LIST_CAP) {
            if (tid == 0) atomicOr(&ctr[ORB_CNT_ERR], ORB_ERR_LIST);
            size = 0;
            break;
        }
        const int prev = size;
        for (int i = tid; i < size; i += OCT_T) S.cc[i][0] = S.cc[i][1] = S.cc[i][2] = S.cc[i][3] = 0;
        __syncthreads();
        for_keys([&](int x, int y, int& nd) {
            const int i = nd;
            if (S.cnt[cur][i] >= 2) atomicAdd(&S.cc[i][quad_of(S.rx[cur][i], S.ry[cur][i], x, y)], 1);
        });
        // Expandable nodes, sorted by (size, creation order), largest first.
        const bool e = tid < size && S.cnt[cur][tid] >= 2;
        int nE;
        const int ex_e = block_scan(e ? 1 : 0, S.scan, &nE);
        // The keys are distinct (they carry the node index), so a key's
        // place in descending order is the number of larger keys; every
        // thread reads the same key at a time (an LDS broadcast).  S.keep
        // holds each node's place (32767: not expandable).
        if (e)
            S.sortk[ex_e] = ((unsigned long long)S.cnt[cur][tid] << 32) |
                            ((unsigned long long)(unsigned)S.seq[cur][tid] << 16) | (unsigned)tid;
        if (tid < size) S.keep[tid] = 32767;
        __syncthreads();
        if (tid < nE) {
            const unsigned long long mine = S.sortk[tid];
            int rank = 0;
            for (int j2 = 0; j2 < nE; j2++) rank += S.sortk[j2] > mine;
            S.best[rank] = mine;                  // S.best is free until the retain step
            S.keep[mine & 0xFFFF] = (short)rank;
        }
        __syncthreads();
        // Gains in processing order and the first split reaching the quota.
        const int j = tid;
        int node_j = -1, nc = 0;
        if (j < nE) {
            node_j = (int)(S.best[j] & 0xFFFF);
            nc = (S.cc[node_j][0] > 0) + (S.cc[node_j][1] > 0) + (S.cc[node_j][2] > 0) + (S.cc[node_j][3] > 0);
        }
        int gsum;
        const int ex_g = block_scan(j < nE ? nc - 1 : 0, S.scan, &gsum);
        if (tid == 0) S.vars[0] = nE - 1;
        __syncthreads();
        if (j < nE && size + ex_g + nc - 1 >= N) atomicMin(&S.vars[0], j);
        __syncthreads();
        const int jstar = S.vars[0];
        const bool proc = j <= jstar && j < nE;
        // Kept list nodes: all but the processed ones (place <= jstar).  One
        // scan for the processed nodes' children (<= 4 * ORB_LIST_CAP < 2^16)
        // and the kept nodes (<= ORB_LIST_CAP).
        static_assert(4 * ORB_LIST_CAP < (1 << 16) && ORB_LIST_CAP <= OCT_T, "octree list packing");
        const bool kp = tid < size && S.keep[tid] > jstar;
        int tot;
        const int ex = block_scan((proc ? nc : 0) | ((kp ? 1 : 0) << 16), S.scan, &tot);
        const int ex_nc = ex & 0xFFFF, ex_kp = ex >> 16, Cp = tot & 0xFFFF, nkeep = tot >> 16;
        const int nsize = Cp + nkeep;
        if (nsize > ORB_LIST_CAP) {
            if (tid == 0) atomicOr(&ctr[ORB_CNT_ERR], ORB_ERR_LIST);
            size = 0;
            break;
        }
        const int nb = cur ^ 1;
        if (proc) {
            const int i = node_j;
            const int base = Cp - ex_nc - nc;
            int r = 0;
            for (int q = 3; q >= 0; q--) {
                const int c = S.cc[i][q];
                if (c > 0) {
                    const int pos = base + r++;
                    int crx, cry;
                    child_rect(S.rx[cur][i], S.ry[cur][i], q, &crx, &cry);
                    S.rx[nb][pos] = crx;
                    S.ry[nb][pos] = cry;
                    S.cnt[nb][pos] = c;
                    S.seq[nb][pos] = ex_nc + (nc - r);
                    S.map[i][q] = (short)pos;
                } else {
                    S.map[i][q] = -1;
                }
            }
        }
        int keeppos = -1;
        if (kp) {
            const int i = tid;
            keeppos = Cp + ex_kp;
            S.rx[nb][keeppos] = S.rx[cur][i];
            S.ry[nb][keeppos] = S.ry[cur][i];
            S.cnt[nb][keeppos] = S.cnt[cur][i];
            S.seq[nb][keeppos] = S.seq[cur][i];
        }
        __syncthreads();
        if (tid < size) S.keep[tid] = (short)(kp ? keeppos : -1);
        __syncthreads();
        for_keys([&](int x, int y, int& nd) {
            const int n0 = nd;
            const int kv = S.keep[n0];
            nd = kv < 0 ? S.map[n0][quad_of(S.rx[cur][n0], S.ry[cur][n0], x, y)] : kv;
        });
        cur ^= 1;
        size = nsize;
        __syncthreads();
        if (size >= N || size == prev) break;
    }

    // ---- retain the best response per node; first in key order on ties
    for (int i = tid; i < size; i += OCT_T) S.best[i] = 0ull;
    __syncthreads();
    auto retain = [&](unsigned long long key, int nd) {
        const unsigned long long packed = (key & 0xFFFFFFFF00000000ull) | (0xFFFFFFFFull - (key & 0xFFFFFFFFull));
        atomicMax(&S.best[nd], packed);
    };
#pragma unroll
    for (int m = 0; m < OCT_REG; m++)
        if (tid + OCT_T * m < K) retain(gload8u64(keys + tid + OCT_T * m), kn[m]);
    for (int k = tid + OCT_T * OCT_REG; k < K; k += OCT_T) retain(keys[k], node[k]);
    __syncthreads();
    const int cap = min(L.list_cap, ORB_LIST_CAP);
    if (size > cap) {
        if (tid == 0) atomicOr(&ctr[ORB_CNT_ERR], ORB_ERR_LIST);
        size = 0;
    }
    for (int i = tid; i < size; i += OCT_T) {
        const unsigned long long p = S.best[i];
        gstore8(L.ret + i, (p & 0xFFFFFFFF00000000ull) | (0xFFFFFFFFull - (p & 0xFFFFFFFFull)));
    }
    if (tid == 0) ctr[ORB_CNT_RET + level] = size;
}

// ============================================================== describe
// fastAtan2 (OpenCV 2.4/3.x): degree-7 odd polynomial, degrees in [0,360).
__device__ __forceinline__ float fast_atan2_dev(float y, float x)
{
    const float P1 = 0.9997878412794807f * (float)(180 / 3.1415926535897932384626433832795);
    const float P3 = -0.3258083974640975f * (float)(180 / 3.1415926535897932384626433832795);
    const float P5 = 0.1555786518463281f * (float)(180 / 3.1415926535897932384626433832795);
    const float P7 = -0.04432655554792128f * (float)(180 / 3.1415926535897932384626433832795);
    const float ax = fabsf(x), ay = fabsf(y);
    float a, c, c2;
    if (ax >= ay) {
        c = __fdiv_rn(ay, ax + (float)DBL_EPSILON);
        c2 = c * c;
        a = (((P7 * c2 + P5) * c2 + P3) * c2 + P1) * c;
    } else {
        c = __fdiv_rn(ax, ay + (float)DBL_EPSILON);
        c2 = c * c;
        a = 90.f - (((P7 * c2 + P5) * c2 + P3) * c2 + P1) * c;
    }
    if (x < 0) a = 180.f - a;
    if (y < 0) a = 360.f - a;
    return a;
}

// cos/sin of a float angle, evaluated in double by Cody-Waite reduction and
// Taylor series, then rounded to float: a fixed method so that the CPU
// statement and this kernel agree bit for bit.
__device__ __forceinline__ void sincos_canonical_dev(double x, double* s, double* c)
{
    const double kPio2Hi = 1.57079632673412561417e+00;
    const double kPio2Lo = 6.07710050650619224932e-11;
    const double kTwoOverPi = 6.36619772367581382433e-01;
    const double kd = rint(x * kTwoOverPi);
    const int k = (int)kd;
    const double r = (x - kd * kPio2Hi) - kd * kPio2Lo;
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;
    ps = ps * r2 - 1.0 / 1307674368000.0;
    ps = ps * r2 + 1.0 / 6227020800.0;
    ps = ps * r2 - 1.0 / 39916800.0;
    ps = ps * r2 + 1.0 / 362880.0;
    ps = ps * r2 - 1.0 / 5040.0;
    ps = ps * r2 + 1.0 / 120.0;
    ps = ps * r2 - 1.0 / 6.0;
    const double sr = r + r * r2 * ps;
    double pc = 1.0 / 6402373705728000.0;
    pc = pc * r2 - 1.0 / 20922789888000.0;
    pc = pc * r2 + 1.0 / 87178291200.0;
    pc = pc * r2 - 1.0 / 479001600.0;
    pc = pc * r2 + 1.0 / 3628800.0;
    pc = pc * r2 - 1.0 / 40320.0;
    pc = pc * r2 + 1.0 / 720.0;
    pc = pc * r2 - 1.0 / 24.0;
    pc = pc * r2 + 0.5;
    const double cr = 1.0 - r2 * pc;
    switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

// A 16-lane group per kept keypoint, four per wavefront; a workgroup serves
// 16 keypoints of one (image, level), so the level's geometry and output
// offset are uniform across it.  Lane j of a group holds patch rows j, j+16
// (and j+32), each loaded with constant column offsets, and BRIEF pairs
// 16j .. 16j+15 (descriptor bytes 2j, 2j+1).
#define DESC_G 16
static_assert(256 / DESC_G == ORB_DESC_KPB, "describe blocks are counted on the host");
__global__ __launch_bounds__(256) void describe_kernel(const ImageJob* __restrict__ jobs,
                                                       const RunTable* __restrict__ rt)
{
    __shared__ uint2 s_brief[ORB_DESC_KPB][37 * 5];   // per group: 37 rows x 40 bytes
    __shared__ uint4 s_pat[64];
    const int sub = threadIdx.x & (DESC_G - 1), grp = threadIdx.x / DESC_G;
    const int nseg = rt->njobs * rt->max_levels;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int dpj = rt->desc_per_job;
    const int f = dpj > 0 ? (blk / dpj) * rt->max_levels +
                                find_segment(rt->desc_start + (blk / dpj) * rt->max_levels, rt->max_levels, blk)
                          : find_segment(rt->desc_start, nseg, blk);
    const int jb = f / rt->max_levels, level = f % rt->max_levels;
    const ImageJob& J = jobs[jb];
    const int* ctr = J.counters;
    int base = 0, total = 0;
    for (int l = 0; l < J.nlevels; l++) {
        const int c = (int)gload4((const uint8_t*)(ctr + ORB_CNT_RET + l));
        base += l < level ? c : 0;
        total += c;
    }
    const int idx = (blk - rt->desc_start[f]) * ORB_DESC_KPB + grp;
    if (level == 0 && idx == 0 && sub == 0) J.counters[ORB_CNT_NOUT] = total;
    // The 256 BRIEF pairs (one dword of 4 signed bytes each) into LDS, laid
    // out so that lane j's k-th 16-byte read holds its pairs 16j+4k .. +3
    // and the 16 lanes of a group read consecutive 16-byte words.
    {
        const int t = threadIdx.x;
        ((unsigned*)s_pat)[(((t >> 2) & 3) * 16 + (t >> 4)) * 4 + (t & 3)] = ((const unsigned*)c_pattern)[t];
    }
    __syncthreads();
    if (level >= J.nlevels || idx >= (int)gload4((const uint8_t*)(ctr + ORB_CNT_RET + level))) return;   // the whole group leaves
    const LevelJob& L = J.lv[level];
    const unsigned long long key = gload8u64(L.ret + idx);
    int xr, yr;
    key_xy(key, L.wCell, L.hCell, &xr, &yr);
    const int cx = xr + ORB_MIN_BORDER, cy = yr + ORB_MIN_BORDER;
    const int pitch = L.pitch;

    // Both patches are read with dword-aligned loads issued together, before
    // any use: the 31x31 orientation patch of the level and the 37x37 window
    // of the blurred level that every rotated BRIEF sample falls in (pattern
    // radius <= 18.39 px).  Keypoints lie >= 19 px inside the level (x <= w-20,
    // y <= h-20), so every dword lies in rows that have a successor and ends
    // at most 2 bytes past the row's last pixel: inside the level's buffer.
    // Blur window rows past the patch (37..47) are clamped to its last row
    // and unused.
    const int xa0 = (cx - 15) & ~3;               // 31 px + <= 3 of alignment: 9 dwords a row
    const int xb0 = (cx - 18) & ~3;               // 37 px + <= 3 of alignment: 10 dwords a row
    // A row is read as two 16-byte loads and the rest.
    unsigned pa[2][9], pb[3][10];
    // Disc rows: lane j holds v = -j and v = +j, which share one disc mask.
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint8_t* rp = L.lvl + (size_t)(cy + (h ? sub : -sub)) * pitch + xa0;
        const uint4 q0 = gload16(rp), q1 = gload16(rp + 16);
        pa[h][0] = q0.x, pa[h][1] = q0.y, pa[h][2] = q0.z, pa[h][3] = q0.w;
        pa[h][4] = q1.x, pa[h][5] = q1.y, pa[h][6] = q1.z, pa[h][7] = q1.w;
        pa[h][8] = gload4(rp + 32);
    }
#pragma unroll
    for (int h = 0; h < 3; h++) {
        const uint8_t* rp = L.blur + (size_t)(cy - 18 + min(sub + 16 * h, 36)) * pitch + xb0;
        const uint4 q0 = gload16(rp), q1 = gload16(rp + 16);
        const uint2 q2 = gload8(rp + 32);
        pb[h][0] = q0.x, pb[h][1] = q0.y, pb[h][2] = q0.z, pb[h][3] = q0.w;
        pb[h][4] = q1.x, pb[h][5] = q1.y, pb[h][6] = q1.z, pb[h][7] = q1.w;
        pb[h][8] = q2.x, pb[h][9] = q2.y;
    }

    // IC_Angle (ORBextractor.cc:94-141): m10 = sum u*I, m01 = sum v*I over
    // the disc |u| <= umax[|v|].  Per dword of 4 pixels: the disc's byte mask
    // (a table over the row's |v| and the patch's alignment), then
    // v_dot4_u32_u8 for the masked sum and for sum (u+20)*I.
    const int al = (cx - 15) & 3;                 // u of byte k of dword c: 4c + k - 15 - al
    const unsigned w0 = (unsigned)(5 - al) * 0x01010101u + 0x03020100u;   // bytes u+20 of dword 0
    int m10 = 0, m01 = 0;
    const unsigned* mk = c_icmask + (al * 16 + sub) * 9;
    unsigned mask[9];
#pragma unroll
    for (int c = 0; c < 9; c++) mask[c] = mk[c];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int v = h ? sub : -sub;
        unsigned s1 = 0, su = 0;
#pragma unroll
        for (int c = 0; c < 9; c++) {
            const unsigned md = pa[h][c] & mask[c];
            s1 = __builtin_amdgcn_udot4(md, 0x01010101u, s1, false);
            su = __builtin_amdgcn_udot4(md, w0 + (unsigned)(4 * c) * 0x01010101u, su, false);
        }
        if (h == 0 || sub > 0) {                  // row v = 0 once
            m01 += v * (int)s1;
            m10 += (int)su - 20 * (int)s1;
        }
    }
    m10 = (int)row16_sum((unsigned)m10);
    m01 = (int)row16_sum((unsigned)m01);
    const float angle_deg = fast_atan2_dev((float)m01, (float)m10);

    const float factorPI = (float)(3.1415926535897932384626433832795 / 180.f);
    const float angle = angle_deg * factorPI;
    double sd, cd;
    sincos_canonical_dev((double)angle, &sd, &cd);
    const float a = (float)cd, b = (float)sd;

    // The blur window goes to LDS only now, so that its reads stay in flight
    // under the orientation and rotation arithmetic above.
    uint2* win = s_brief[grp];
#pragma unroll
    for (int h = 0; h < 3; h++) {
        const int r = sub + 16 * h;
        if (r < 37) {
#pragma unroll
            for (int c = 0; c < 5; c++) win[r * 5 + c] = make_uint2(pb[h][2 * c], pb[h][2 * c + 1]);
        }
    }

    // Rotated BRIEF (ORBextractor.cc:153-204) on the blurred level, sampled
    // from the staged window (row stride 40 bytes, centre at row 18).
    // cvRound of a sample coordinate t (|t| < 19) is read from the bits of
    // t + 1.5 * 2^23, whose low 23 bits hold 2^22 + rint(t): the sum rounds
    // to the nearest integer, ties to even, as rint does.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float kRound = 12582912.0f;
    // sample (iy, ix) is byte (18 + iy) * 40 + (cx - xb0) + ix of the window;
    // with both read as 2^22 + i in the low 24 bits, the byte is
    // (2^22 + iy) * 40 + (bits of x) - kBias
    const unsigned kBias = (1u << 22) * 40u + 0x4B400000u - 18u * 40u - (unsigned)(cx - xb0);
    const uint8_t* wb = (const uint8_t*)win;
    unsigned bits = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint4 q = s_pat[k * 16 + sub];
        const unsigned qd[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int d = (int)qd[e];
            const float px0 = (float)(signed char)d, py0 = (float)(signed char)(d >> 8);
            const float px1 = (float)(signed char)(d >> 16), py1 = (float)(d >> 24);
            const float y0r = px0 * b + py0 * a, x0r = px0 * a - py0 * b;
            const float y1r = px1 * b + py1 * a, x1r = px1 * a - py1 * b;
            const unsigned o0 = (__float_as_uint(y0r + kRound) & 0xFFFFFFu) * 40u + __float_as_uint(x0r + kRound) - kBias;
            const unsigned o1 = (__float_as_uint(y1r + kRound) & 0xFFFFFFu) * 40u + __float_as_uint(x1r + kRound) - kBias;
            const int i0 = wb[o0], i1 = wb[o1];
            bits |= (unsigned)(i0 < i1) << (4 * k + e);
        }
    }
    const int out = base + idx;
    gstore2(J.desc + (size_t)out * 32 + 2 * sub, (unsigned short)bits);
    if (sub == 0) {
        const float sc = L.scale;
        OrbKeyPoint kp;
        float fx = (float)cx, fy = (float)cy;
        if (level != 0) {
            fx = fx * sc;
            fy = fy * sc;
        }
        kp.x = fx;
        kp.y = fy;
        kp.size = (float)(int)(31 * sc);
        kp.angle = angle_deg;
        kp.response = (float)(unsigned)(key >> 32);
        kp.octave = level;
        kp.class_id = -1;
        uint8_t* kd = (uint8_t*)(J.kps + out);     // 28 bytes: 16 + 8 + 4
        gstore16(kd, make_uint4(__float_as_uint(kp.x), __float_as_uint(kp.y), __float_as_uint(kp.size),
                                __float_as_uint(kp.angle)));
        *(__attribute__((address_space(1))) orb_v2u*)(kd + 16) =
            orb_v2u{__float_as_uint(kp.response), (unsigned)kp.octave};
        gstore4(kd + 24, (unsigned)kp.class_id);
    }
}

// ============================================================ zero fill
// Counters and per-cell flags of every job (one workgroup per job).
__global__ void zero_counters_kernel(const ImageJob* __restrict__ jobs, int njobs)
{
    const int j = blockIdx.x;
    if (j >= njobs) return;
    const ImageJob& J = jobs[j];
    if (threadIdx.x < ORB_CNT_SLOTS) gstore4((uint8_t*)(J.counters + threadIdx.x), 0u);
    // 32 threads per level, the levels side by side
    for (int l = threadIdx.x >> 5; l < J.nlevels; l += blockDim.x >> 5) {
        uint8_t* f = J.lv[l].cellflag;
        const int n = J.lv[l].ncells;
        for (int i = threadIdx.x & 31; i < n; i += 32) gstore1(f + i, 0);
    }
}

// ========================================================= Hamming helpers
__device__ __forceinline__ int hamming32(const uint4 a0, const uint4 a1, const uint8_t* b)
{
    const uint4 b0 = *(const uint4*)b, b1 = *(const uint4*)(b + 16);
    return __popc(a0.x ^ b0.x) + __popc(a0.y ^ b0.y) + __popc(a0.z ^ b0.z) + __popc(a0.w ^ b0.w) +
           __popc(a1.x ^ b1.x) + __popc(a1.y ^ b1.y) + __popc(a1.z ^ b1.z) + __popc(a1.w ^ b1.w);
}

// ================================================================ stereo
// Frame::ComputeStereoMatches (Frame.cc:831-1082), a 16-lane group per left
// keypoint.  Search: right keypoints whose row band [floor(y-2s), ceil(y+2s)]
// holds row (size_t)vL, octave within +-1, uR in [uL-maxD, uL]; best is the
// smallest (distance, index) below TH_HIGH.  Refinement: 11x11 SAD of
// centre-subtracted patches over 11 shifts, parabola fit, depth = bf/disp.

// Group each pair's right keypoints by (octave, level-0 row bucket)
// (counting sort in LDS, one workgroup per pair) and lay them out in that
// order as 16-byte records {minr, maxr, index << 4 | octave, x} (the row band
// of Frame.cc:867-882, with the same float statements) and 32-byte
// descriptors; plus, per octave and level-0 row, the entry range of that
// octave's buckets that can hold the row in a band.  Order inside a bucket is free: the search keeps the smallest
// (distance, index), which does not depend on the visiting order.
#define STEREO_NB 4096               // buckets: rows >> qs
#define STEREO_T 1024
#define STEREO_PER_T 4               // right keypoints per thread (<= 4096)
__global__ __launch_bounds__(STEREO_T) void stereo_prep_kernel(const StereoArgs* __restrict__ args)
{
    __shared__ int s_cnt[STEREO_NB + 1];
    __shared__ int s_part[STEREO_T / 64];
    const StereoArgs& A = args[blockIdx.x];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int n = min(min(*A.nr, A.cap_r), STEREO_T * STEREO_PER_T);
    const int nrows = A.nrows, nlev = A.nlev;
    int qs = 0;
    while (nlev * (((nrows - 1) >> qs) + 1) > STEREO_NB) qs++;
    const int nb = ((nrows - 1) >> qs) + 1;      // row buckets per octave
    const int nk = nlev * nb;                    // keys: octave * nb + row bucket
    // Bucket of a y: floor(y) clamped to the rows (NaN to row 0), >> qs.
    auto bucket = [nrows, qs](float y) {
        const float f = floorf(y);
        const int r = !(f > 0.f) ? 0 : (f >= (float)(nrows - 1) ? nrows - 1 : (int)f);
        return r >> qs;
    };
    for (int j = t; j <= nk; j += STEREO_T) s_cnt[j] = 0;
    __syncthreads();
    // Each thread's keypoints and descriptors are read once, here; the
    // scatter after the scan uses the registers.
    int eb[STEREO_PER_T], er[STEREO_PER_T], ko[STEREO_PER_T];
    uint2 kxy[STEREO_PER_T];
    uint4 kd0[STEREO_PER_T], kd1[STEREO_PER_T];
#pragma unroll
    for (int k = 0; k < STEREO_PER_T; k++) {
        const int i = t + STEREO_T * k;
        eb[k] = 0;
        ko[k] = 0;
        kxy[k] = make_uint2(0u, 0u);
        kd0[k] = kd1[k] = make_uint4(0u, 0u, 0u, 0u);
        if (i < n) {
            const uint8_t* kb = (const uint8_t*)(A.kr + i);
            kxy[k] = gload8(kb + offsetof(OrbKeyPoint, x));
            ko[k] = (int)gload4(kb + offsetof(OrbKeyPoint, octave));
            const uint8_t* d = A.dr + (size_t)i * 32;
            kd0[k] = gload16(d);
            kd1[k] = gload16(d + 16);
            eb[k] = min(ko[k], nlev - 1) * nb + bucket(__uint_as_float(kxy[k].y));
        }
        er[k] = i < n ? atomicAdd(&s_cnt[eb[k]], 1) : 0;
    }
    __syncthreads();
    // Exclusive scan of the nk counts: 4 per thread, wavefront scans, then
    // the 16 wavefront totals.
    int loc[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int j = 4 * t + k;
        loc[k] = sum;
        sum += j < nk ? s_cnt[j] : 0;
    }
    int inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) s_part[wv] = inc;
    __syncthreads();
    int base = inc - sum;
    for (int k = 0; k < wv; k++) base += s_part[k];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int j = 4 * t + k;
        if (j < nk) s_cnt[j] = base + loc[k];
    }
    if (t == 0) s_cnt[nk] = n;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < STEREO_PER_T; k++) {
        const int i = t + STEREO_T * k;
        if (i < n) {
            const uint2 xy = kxy[k];
            const int oct = ko[k];
            const float ky = __uint_as_float(xy.y);
            const float r = 2.0f * A.scale[oct];
            const int maxr = (int)ceilf(ky + r);
            const int minr = (int)floorf(ky - r);
            const int pos = s_cnt[eb[k]] + er[k];
            gstore16((uint8_t*)(A.srec + pos), make_uint4((unsigned)minr, (unsigned)maxr, (unsigned)((i << 4) | oct), xy.x));
            uint8_t* o = A.sdesc + (size_t)pos * 32;
            gstore16(o, kd0[k]);
            gstore16(o + 16, kd1[k]);
        }
    }
    // A candidate of octave o for row `row` has y in (row - 1 - r, row + 1 + r),
    // r = 2 * scale[o]; one more row each side absorbs the rounding of these
    // bounds.  rows[o * nrows + row]: the entries of those buckets of octave o.
    for (int idx = t; idx < nlev * nrows; idx += STEREO_T) {
        const int o = idx / nrows, row = idx - o * nrows;
        const float r = 2.0f * A.scale[o];
        const int blo = bucket((float)row - 2.0f - r), bhi = bucket((float)row + 2.0f + r);
        *(__attribute__((address_space(1))) orb_v2u*)(A.rows + idx) =
            orb_v2u{(unsigned)s_cnt[o * nb + blo], (unsigned)s_cnt[o * nb + bhi + 1]};
    }
}

// One 32-bit load of a patch row, bytewise where it would pass the level's
// right edge.
__device__ __forceinline__ unsigned patch_dword(const uint8_t* rp, int x, int w)
{
    if (x + 3 < w) return *(const unsigned*)(rp + x);
    unsigned d = 0;
    for (int b = 0; b < 4; b++)
        if (x + b < w) d |= (unsigned)rp[x + b] << (8 * b);
    return d;
}


// grid (ceil(max left capacity / 16), pairs): lanes 16g .. 16g+15 of a
// wavefront serve one left keypoint.  Every branch is uniform within a
// group, and cross-lane steps stay inside it.
//   search: the group's lanes stride over the row's candidate range;
//   refinement: lane r < 11 holds patch row r of both images in registers
//   as 16-bit pairs and accumulates its row's part of the 11 shift SADs with
//   v_sad_u16 (two absolute differences per instruction), as
//   |(l - cL) - (r - cR)| = |(l + cR - cL + 256) - (r + 256)|, both sides in
//   [0, 766]; the row parts are summed across the group with DPP, two shifts
//   per dword (a total is at most 121 * 510 < 2^16).
#define ST_G 16
#ifndef ST_CHUNK
#define ST_CHUNK 2
#endif
__global__ __launch_bounds__(256) void stereo_kernel(const StereoArgs* __restrict__ args)
{
    const int sub = threadIdx.x & (ST_G - 1);
    const int gbase = threadIdx.x & 63 & ~(ST_G - 1);
    // XCD-aware order: each XCD takes a contiguous run of (pair, keypoint
    // block), so the patch rows that a pair's neighbouring keypoints share
    // are fetched into one L2 rather than into all of them.
    const int lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int pair = lb / gridDim.x, kblk = lb - pair * gridDim.x;
    const StereoArgs& A = args[pair];
    const int iL = kblk * (256 / ST_G) + (int)(threadIdx.x / ST_G);
    const int nl = *A.nl, nr = *A.nr;
    if (iL >= nl || iL >= A.cap_l) return;             // the whole group leaves
    float outU = -1.0f, outD = -1.0f;
    int outSad = -1;
    const uint8_t* kpb = (const uint8_t*)(A.kl + iL);
    const uint2 kxy = gload8(kpb + offsetof(OrbKeyPoint, x));
    const uint4 a0 = gload16(A.dl + (unsigned)iL * 32u), a1 = gload16(A.dl + (unsigned)iL * 32u + 16u);
    const int levelL = (int)gload4(kpb + offsetof(OrbKeyPoint, octave));
    const float vL = __uint_as_float(kxy.y), uL = __uint_as_float(kxy.x);
    const int row = (int)vL;
    const float minZ = A.mb;
    const float minD = 0;
    const float maxD = A.bf / minZ;
    const float minU = uL - maxD;
    const float maxU = uL - minD;
    bool go = row >= 0 && row < A.nrows && !(maxU < 0);
    // The level both patches are read from (octaves are validated on entry
    // and produced in range by the extractor).
    const LevelJob& LL = A.left->lv[levelL];
    const LevelJob& LR = A.right->lv[levelL];
    const uint8_t* pl = LL.lvl;
    const uint8_t* pr = LR.lvl;
    const int pitchL = LL.pitch, pitchR = LR.pitch, wL = LL.w, hL = LL.h, wR = LR.w, hR = LR.h;
    int bestDist = 100, bestIdx = INT_MAX;
    float bestU = 0.f;
    if (go) {
        (void)nr;
        // Octaves levelL-1 .. levelL+1 only (Frame.cc:929), each over the
        // entries of its own row band.
        for (int oc = max(levelL - 1, 0); oc <= min(levelL + 1, A.nlev - 1); oc++) {
            const int2 rg = A.rows[oc * A.nrows + row];
            // Chunks of ST_CHUNK candidates per lane: the records are loaded
            // together, then the descriptors of those that pass the band, octave
            // and uR tests, so each chunk waits on memory twice.
            for (int c0 = rg.x + sub; c0 < rg.y; c0 += ST_CHUNK * ST_G) {
                int4 rec[ST_CHUNK];
#pragma unroll
                for (int u = 0; u < ST_CHUNK; u++) {
                    const int c = c0 + u * ST_G;
                    if (c < rg.y) {
                        const uint4 q = gload16((const uint8_t*)A.srec + (unsigned)c * 16u);
                        rec[u] = make_int4((int)q.x, (int)q.y, (int)q.z, (int)q.w);
                    } else {
                        rec[u] = make_int4(1, 0, 0, 0);   // minr > maxr: never a candidate
                    }
                }
                bool pass[ST_CHUNK];
#pragma unroll
                for (int u = 0; u < ST_CHUNK; u++) {
                    const int oct = rec[u].z & 15;
                    const float uR = __int_as_float(rec[u].w);
                    pass[u] = row >= rec[u].x && row <= rec[u].y && oct >= levelL - 1 && oct <= levelL + 1 &&
                              uR >= minU && uR <= maxU;
                }
                uint4 b0[ST_CHUNK], b1[ST_CHUNK];
#pragma unroll
                for (int u = 0; u < ST_CHUNK; u++) {
                    if (pass[u]) {
                        const uint8_t* d = A.sdesc + (unsigned)(c0 + u * ST_G) * 32u;
                        b0[u] = gload16(d);
                        b1[u] = gload16(d + 16);
                    }
                }
#pragma unroll
                for (int u = 0; u < ST_CHUNK; u++) {
                    if (pass[u]) {
                        const int iR = rec[u].z >> 4;
                        const int dist = __popc(a0.x ^ b0[u].x) + __popc(a0.y ^ b0[u].y) + __popc(a0.z ^ b0[u].z) +
                                         __popc(a0.w ^ b0[u].w) + __popc(a1.x ^ b1[u].x) + __popc(a1.y ^ b1[u].y) +
                                         __popc(a1.z ^ b1[u].z) + __popc(a1.w ^ b1[u].w);
                        if (dist < bestDist || (dist == bestDist && iR < bestIdx)) {
                            bestDist = dist;
                            bestIdx = iR;
                            bestU = __int_as_float(rec[u].w);
                        }
                    }
                }
            }
        }
        // Smallest (distance, index) across the group.
#pragma unroll
        for (int o = ST_G / 2; o > 0; o >>= 1) {
            const int od = __shfl_xor(bestDist, o, ST_G), oi = __shfl_xor(bestIdx, o, ST_G);
            const float ou = __shfl_xor(bestU, o, ST_G);
            if (od < bestDist || (od == bestDist && oi < bestIdx)) {
                bestDist = od;
                bestIdx = oi;
                bestU = ou;
            }
        }
        go = bestDist < 75 && bestIdx != INT_MAX;
    }
    int sx0 = 0, sy0 = 0, sxR = 0;
    float scaleduR0 = 0.f;
    if (go) {
        const float uR0 = bestU;
        const float scaleFactor = A.inv_scale[levelL];
        const float scaleduL = roundf(uL * scaleFactor);
        const float scaledvL = roundf(vL * scaleFactor);
        scaleduR0 = roundf(uR0 * scaleFactor);
        const float iniu = scaleduR0 - 5 - 5;
        const float endu = scaleduR0 + 5 + 5 + 1;
        if (iniu < 0 || endu >= wR) go = false;
        sx0 = (int)scaleduL;
        sy0 = (int)scaledvL;
        sxR = (int)scaleduR0;
        // Keypoints of this extractor always leave room for the 11x11
        // patches; caller-supplied ones are checked instead of read past the
        // level (the reference would index outside the cv::Mat).
        if (sy0 - 5 < 0 || sy0 + 5 >= hL || sy0 + 5 >= hR || sx0 - 5 < 0 || sx0 + 5 >= wL) go = false;
    }
    if (go) {
        // Lane r < 11: patch row r, left bytes l0..l10 (x = sx0-5 ..) and
        // right bytes r0..r20 (x = sxR-10 ..); lanes 11..15 repeat row 10
        // and contribute nothing.
        const int r = sub < 11 ? sub : 10;
        const uint8_t* rl = pl + (unsigned)((sy0 - 5 + r) * pitchL);
        const uint8_t* rr = pr + (unsigned)((sy0 - 5 + r) * pitchR);
        const int xl0 = (sx0 - 5) & ~3, offL = sx0 - 5 - xl0;
        const int xr0 = (sxR - 10) & ~3, offR = sxR - 10 - xr0;
        unsigned dl[4], dr[6];
        if (xl0 + 15 < wL) {                     // the whole 16 bytes lie in the row
            const uint4 q = gload16(rl + xl0);
            dl[0] = q.x, dl[1] = q.y, dl[2] = q.z, dl[3] = q.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) dl[k] = patch_dword(rl, xl0 + 4 * k, wL);
        }
        if (xr0 + 23 < wR) {
            const uint4 q = gload16(rr + xr0);
            const uint2 q2 = gload8(rr + xr0 + 16);
            dr[0] = q.x, dr[1] = q.y, dr[2] = q.z, dr[3] = q.w, dr[4] = q2.x, dr[5] = q2.y;
        } else {
#pragma unroll
            for (int k = 0; k < 6; k++) dr[k] = patch_dword(rr, xr0 + 4 * k, wR);
        }
        unsigned la[3], ra[6];
#pragma unroll
        for (int k = 0; k < 3; k++) la[k] = __builtin_amdgcn_alignbyte(dl[k + 1], dl[k], (unsigned)offL);
#pragma unroll
        for (int k = 0; k < 5; k++) ra[k] = __builtin_amdgcn_alignbyte(dr[k + 1], dr[k], (unsigned)offR);
        ra[5] = dr[5] >> (8 * offR);
        // Centres from row 5: cL = l5, cR of shift s = r(s+5), s = 0..10.
        const int cL = __shfl((int)((la[1] >> 8) & 0xFFu), gbase + 5, 64);
        const unsigned c0 = (unsigned)__shfl((int)__builtin_amdgcn_alignbyte(ra[2], ra[1], 1u), gbase + 5, 64);
        const unsigned c1 = (unsigned)__shfl((int)__builtin_amdgcn_alignbyte(ra[3], ra[2], 1u), gbase + 5, 64);
        const unsigned c2 = (unsigned)__shfl((int)__builtin_amdgcn_alignbyte(ra[4], ra[3], 1u), gbase + 5, 64);
        // Left pairs (l2k, l2k+1), k < 5, and (l10, 0).
        unsigned lp[6];
#pragma unroll
        for (int k = 0; k < 5; k++) lp[k] = __builtin_amdgcn_perm(0u, la[k >> 1], (k & 1) ? 0x0C030C02u : 0x0C010C00u);
        lp[5] = (la[2] >> 16) & 0xFFu;
        // Right pairs (rj, rj+1) + 256 each, j < 20, and (rj, 0) + 256, j = 10..20.
        unsigned rp[20], rs[11];
#pragma unroll
        for (int j = 0; j < 20; j++) {
            const int q = j >> 2, p = j & 3;
            rp[j] = __builtin_amdgcn_perm(q + 1 < 6 ? ra[q + 1] : 0u, ra[q],
                                          0x0C000C00u | ((unsigned)(p + 1) << 16) | (unsigned)p) |
                    0x01000100u;
        }
#pragma unroll
        for (int j = 10; j <= 20; j++) rs[j - 10] = ((ra[j >> 2] >> (8 * (j & 3))) & 0xFFu) | 0x100u;
        unsigned part[11];
#pragma unroll
        for (int s = 0; s < 11; s++) {
            const unsigned cw = s < 4 ? c0 : (s < 8 ? c1 : c2);
            const unsigned e = ((cw >> (8 * (s & 3))) & 0xFFu) + 256u - (unsigned)cL;   // cR - cL + 256
            const unsigned e2 = e | (e << 16);
            unsigned acc = 0;
#pragma unroll
            for (int k = 0; k < 5; k++) acc = __builtin_amdgcn_sad_u16(lp[k] + e2, rp[s + 2 * k], acc);
            part[s] = __builtin_amdgcn_sad_u16(lp[5] + e, rs[s], acc);
        }
        unsigned tot[6];
#pragma unroll
        for (int m = 0; m < 6; m++) {
            const unsigned v = 2 * m + 1 < 11 ? (part[2 * m] | (part[2 * m + 1] << 16)) : part[2 * m];
            tot[m] = row16_sum(sub < 11 ? v : 0u);
        }
        int sad[11];
#pragma unroll
        for (int s = 0; s < 11; s++) sad[s] = (int)((tot[s >> 1] >> (16 * (s & 1))) & 0xFFFFu);
        // First minimum over inc = -5..5, then the parabola on its
        // neighbours, evaluated identically by every lane of the group.
        unsigned key = 0xFFFFFFFFu;
#pragma unroll
        for (int s = 0; s < 11; s++) key = min(key, ((unsigned)sad[s] << 4) | (unsigned)s);
        const int best = (int)(key >> 4), bs = (int)(key & 15u);
        const int bestinc = bs - 5;
        int d1 = 0, d3 = 0;
#pragma unroll
        for (int s = 0; s < 11; s++) {
            d1 = s == bs - 1 ? sad[s] : d1;
            d3 = s == bs + 1 ? sad[s] : d3;
        }
        if (bestinc != -5 && bestinc != 5) {
            const float dist1 = (float)d1;
            const float dist2 = (float)best;
            const float dist3 = (float)d3;
            const float deltaR = __fdiv_rn(dist1 - dist3, 2.0f * (dist1 + dist3 - 2.0f * dist2));
            if (!(deltaR < -1 || deltaR > 1)) {
                float bestuR = A.scale[levelL] * ((float)scaleduR0 + (float)bestinc + deltaR);
                float disparity = (uL - bestuR);
                if (disparity >= minD && disparity < maxD) {
                    if (disparity <= 0) {
                        disparity = 0.01f;
                        bestuR = (float)((double)uL - 0.01);
                    }
                    outD = __fdiv_rn(A.bf, disparity);
                    outU = bestuR;
                    outSad = best;
                }
            }
        }
    }
    if (sub == 0) {
        A.uRight[iL] = outU;
        A.depth[iL] = outD;
        A.match_sad[iL] = outSad;
    }
}

// Median cut (Frame.cc:1066-1081): matches whose SAD is not below
// 1.5 * 1.4 * median are dropped, the median being element M/2 of the M
// kept SADs in ascending order.  One workgroup per pair selects it with two
// 256-bin histograms (SADs are below 2^16), high byte then low byte, and
// clears the rejected keypoints.
#define MEDIAN_T 1024
#define MEDIAN_PER_T 4               // left keypoints per thread (<= 4096)

// Wavefront 0: the bin of s_h holding rank k (0-based) of the ascending
// order, and k's offset inside it; bin -1 when the histogram holds <= k.
__device__ __forceinline__ void median_pick(const int* s_h, int k, int* s_out)
{
    const int lane = threadIdx.x;
    int c[4], s = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        c[j] = s_h[4 * lane + j];
        s += c[j];
    }
    int inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    int ex = inc - s;
    if (lane == 63 && inc <= k) s_out[0] = -1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (ex <= k && k < ex + c[j]) {
            s_out[0] = 4 * lane + j;
            s_out[1] = k - ex;
        }
        ex += c[j];
    }
}

__global__ __launch_bounds__(MEDIAN_T) void stereo_median_kernel(const StereoArgs* __restrict__ args)
{
    __shared__ int s_h[256];
    __shared__ int s_m;
    __shared__ int s_sel[2];
    const StereoArgs& A = args[blockIdx.x];
    const int t = threadIdx.x;
    const int n = min(min(*A.nl, A.cap_l), MEDIAN_T * MEDIAN_PER_T);
    if (t < 256) s_h[t] = 0;
    if (t == 0) s_m = 0;
    __syncthreads();
    int sad[MEDIAN_PER_T], m = 0;
#pragma unroll
    for (int k = 0; k < MEDIAN_PER_T; k++) {
        const int i = t + MEDIAN_T * k;
        sad[k] = i < n ? (int)gload4((const uint8_t*)(A.match_sad + i)) : -1;
        if (sad[k] >= 0) {
            atomicAdd(&s_h[sad[k] >> 8], 1);
            m++;
        }
    }
    if (m) atomicAdd(&s_m, m);
    __syncthreads();
    const int M = s_m;
    if (M == 0) return;                              // nothing to cut (uniform)
    if (t < 64) median_pick(s_h, M / 2, s_sel);
    __syncthreads();
    const int hb = s_sel[0], k2 = s_sel[1];
    __syncthreads();
    if (t < 256) s_h[t] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MEDIAN_PER_T; k++)
        if (sad[k] >= 0 && (sad[k] >> 8) == hb) atomicAdd(&s_h[sad[k] & 255], 1);
    __syncthreads();
    if (t < 64) median_pick(s_h, k2, s_sel);
    __syncthreads();
    const float median = (float)((hb << 8) | s_sel[0]);
    const float kf = 1.5f * 1.4f;
    const float thDist = kf * median;
#pragma unroll
    for (int k = 0; k < MEDIAN_PER_T; k++) {
        const int i = t + MEDIAN_T * k;
        if (sad[k] >= 0 && !((float)sad[k] < thDist)) {
            gstore4((uint8_t*)(A.uRight + i), __float_as_uint(-1.0f));
            gstore4((uint8_t*)(A.depth + i), __float_as_uint(-1.0f));
        }
    }
}

// ============================================================ brute force
// Best/second-best Hamming match per query with the update rule of
// ORBmatcher::SearchForInitialization (ORBmatcher.cc:614-647): the best is
// the first minimum in index order, the second is the next-smallest distance
// in the multiset; accept iff best <= th_low and best < (float)second*ratio.
__global__ __launch_bounds__(256) void match_bf_kernel(const uint8_t* __restrict__ q, const int* __restrict__ nq_ptr,
                                                       const uint8_t* __restrict__ t, const int* __restrict__ nt_ptr,
                                                       int th_low, float ratio, int* __restrict__ idx,
                                                       int* __restrict__ dist, int cap_q)
{
    const int lane = threadIdx.x & 63;
    const int iq = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nq = *nq_ptr, nt = *nt_ptr;
    if (iq >= nq || iq >= cap_q) return;
    const uint4 a0 = *(const uint4*)(q + (size_t)iq * 32);
    const uint4 a1 = *(const uint4*)(q + (size_t)iq * 32 + 16);
    int d0 = INT_MAX, i0 = INT_MAX, d1 = INT_MAX;
    for (int j = lane; j < nt; j += 64) {
        const int d = hamming32(a0, a1, t + (size_t)j * 32);
        if (d < d0) {
            d1 = d0;
            d0 = d;
            i0 = j;
        } else if (d < d1) {
            d1 = d;
        }
    }
    int bd = d0, bi = i0;
    for (int o = 32; o > 0; o >>= 1) {
        const int od = __shfl_xor(bd, o, 64), oi = __shfl_xor(bi, o, 64);
        if (od < bd || (od == bd && oi < bi)) {
            bd = od;
            bi = oi;
        }
    }
    int sec = (i0 == bi) ? d1 : d0;
    for (int o = 32; o > 0; o >>= 1) sec = min(sec, __shfl_xor(sec, o, 64));
    if (lane == 0) {
        const bool ok = bi != INT_MAX && bd <= th_low && (float)bd < (float)sec * ratio;
        idx[iq] = ok ? bi : -1;
        dist[iq] = bd;
    }
}

// Many (query set, train set) matches in one launch, one wavefront per query.
__device__ __forceinline__ int find_match(const MatchArgs* args, int n, int x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (args[mid].wave_base <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void match_pairs_kernel(const MatchArgs* __restrict__ args, int npairs, int nwaves)
{
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= nwaves) return;
    const MatchArgs& A = args[find_match(args, npairs, wid)];
    const int iq = wid - A.wave_base;
    const int nq = *A.nq, nt = *A.nt;
    if (iq >= nq || iq >= A.cap_q) return;
    const uint4 a0 = *(const uint4*)(A.q + (size_t)iq * 32);
    const uint4 a1 = *(const uint4*)(A.q + (size_t)iq * 32 + 16);
    int d0 = INT_MAX, i0 = INT_MAX, d1 = INT_MAX;
    for (int j = lane; j < nt; j += 64) {
        const int d = hamming32(a0, a1, A.t + (size_t)j * 32);
        if (d < d0) {
            d1 = d0;
            d0 = d;
            i0 = j;
        } else if (d < d1) {
            d1 = d;
        }
    }
    int bd = d0, bi = i0;
    for (int o = 32; o > 0; o >>= 1) {
        const int od = __shfl_xor(bd, o, 64), oi = __shfl_xor(bi, o, 64);
        if (od < bd || (od == bd && oi < bi)) {
            bd = od;
            bi = oi;
        }
    }
    int sec = (i0 == bi) ? d1 : d0;
    for (int o = 32; o > 0; o >>= 1) sec = min(sec, __shfl_xor(sec, o, 64));
    if (lane == 0) {
        const bool ok = bi != INT_MAX && bd <= A.th_low && (float)bd < (float)sec * A.ratio;
        A.idx[iq] = ok ? bi : -1;
        A.dist[iq] = bd;
    }
}

// After the matches of a launch: each pair's query rows become the resident
// "previous" set of its stream (one workgroup per pair).
__global__ __launch_bounds__(256) void save_prev_kernel(const MatchArgs* __restrict__ args, int npairs)
{
    const MatchArgs& A = args[blockIdx.x];
    if (!A.save_desc) return;
    const int n = min(*A.nq, A.save_cap);
    const uint4* src = (const uint4*)A.q;
    uint4* dst = (uint4*)A.save_desc;
    for (int i = threadIdx.x; i < 2 * n; i += 256) dst[i] = src[i];
    if (threadIdx.x == 0) *A.save_n = n;
}

// ============================================================= launchers
// ev (optional): events recorded after each stage -- [0] start, then one per
// pyramid level (max_levels), octree, describe.
extern "C" hipError_t orb_launch_extract(const ImageJob* d_jobs, const RunTable* d_rt, const RunTable* h_rt,
                                         hipStream_t st, hipEvent_t* ev)
{
    int e = 0;
    if (ev) (void)hipEventRecord(ev[e++], st);
    zero_counters_kernel<<<h_rt->njobs, 256, 0, st>>>(d_jobs, h_rt->njobs);
    for (int l = 0; l < h_rt->max_levels; l++) {
        const int nb = h_rt->tile_start[l][h_rt->njobs];
        if (nb > 0) {
            if (h_rt->tile_h[l] == 16) level_kernel<16><<<nb, 256, 0, st>>>(d_jobs, d_rt, l);
            else level_kernel<32><<<nb, 256, 0, st>>>(d_jobs, d_rt, l);
        }
        if (ev) (void)hipEventRecord(ev[e++], st);
    }
    octree_kernel<<<h_rt->njobs * h_rt->max_levels, OCT_T, 0, st>>>(d_jobs, h_rt->max_levels);
    if (ev) (void)hipEventRecord(ev[e++], st);
    const int nblocks = h_rt->desc_start[h_rt->njobs * h_rt->max_levels];
    if (nblocks > 0) describe_kernel<<<nblocks, 256, 0, st>>>(d_jobs, d_rt);
    if (ev) (void)hipEventRecord(ev[e++], st);
    return hipGetLastError();
}

// d_args: npairs StereoArgs in device memory; max_left: the largest cap_l.
// ev (optional): events after the match kernel and after the median cut.
extern "C" hipError_t orb_launch_stereo(const StereoArgs* d_args, int npairs, int max_left, hipStream_t st,
                                        hipEvent_t* ev)
{
    if (npairs <= 0) return hipSuccess;
    stereo_prep_kernel<<<npairs, 1024, 0, st>>>(d_args);
    if (max_left > 0) stereo_kernel<<<dim3((max_left + 256 / ST_G - 1) / (256 / ST_G), npairs), 256, 0, st>>>(d_args);
    if (ev) (void)hipEventRecord(ev[0], st);
    stereo_median_kernel<<<npairs, 1024, 0, st>>>(d_args);
    if (ev) (void)hipEventRecord(ev[1], st);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_match_bf(const uint8_t* q, const int* nq, const uint8_t* t, const int* nt,
                                          int th_low, float ratio, int* idx, int* dist, int cap_q,
                                          hipStream_t st)
{
    if (cap_q > 0) match_bf_kernel<<<(cap_q + 3) / 4, 256, 0, st>>>(q, nq, t, nt, th_low, ratio, idx, dist, cap_q);
    return hipGetLastError();
}

// ============================================================ dense k=2
// All-pairs Hamming search of nq queries against nm map rows, k = 2 nearest
// by (distance, index), the order of ORBmatcher's best/second-best update
// (ORBmatcher.cc:614-647).  The map is split into chunks of KNN_CHUNK rows;
// each (chunk, query tile) workgroup keeps per query the two smallest keys
//   key = distance << 20 | row-in-chunk
// (lexicographic (distance, index) is numeric key order, so the update is
// k1 = med3(k0, k1, key); k0 = min(k0, key)), and a merge kernel reduces the
// chunks.  Two kernels produce the same partial keys:
//   knn_popc_kernel  XOR + v_bcnt popcount, one query per lane, map rows
//                    broadcast from LDS
//   knn_mfma_kernel  bits expanded to int8 +-1; v_mfma_i32_16x16x64_i8 gives
//                    dot = 256 - 2*Hamming exactly (int32 accumulate)
#define KNN_CHUNK 8192
#define KNN_NONE 0xFFFFFFFFu

// Median of three (the compiler emits v_med3_u32 / v_med3_i32).
__device__ __forceinline__ unsigned med3u(unsigned a, unsigned b, unsigned c)
{
    return max(min(a, b), min(max(a, b), c));
}

__device__ __forceinline__ int med3i(int a, int b, int c)
{
    return max(min(a, b), min(max(a, b), c));
}

__global__ __launch_bounds__(256) void knn_popc_kernel(const uint8_t* __restrict__ q, int nq,
                                                       const uint8_t* __restrict__ m, long long nm,
                                                       uint2* __restrict__ part, int nchunks)
{
    __shared__ uint4 rows[256][2];
    const int tid = threadIdx.x;
    const int chunk = blockIdx.x;
    const int qi = blockIdx.y * 256 + tid;
    uint4 a0 = make_uint4(0, 0, 0, 0), a1 = a0;
    if (qi < nq) {
        a0 = *(const uint4*)(q + (size_t)qi * 32);
        a1 = *(const uint4*)(q + (size_t)qi * 32 + 16);
    }
    const long long c0 = (long long)chunk * KNN_CHUNK;
    const int len = (int)min((long long)KNN_CHUNK, nm - c0);
    unsigned k0 = KNN_NONE, k1 = KNN_NONE;
    for (int t0 = 0; t0 < len; t0 += 256) {
        const int n = min(256, len - t0);
        __syncthreads();
        if (tid < n) {
            const uint8_t* r = m + (size_t)(c0 + t0 + tid) * 32;
            rows[tid][0] = *(const uint4*)r;
            rows[tid][1] = *(const uint4*)(r + 16);
        }
        __syncthreads();
        for (int r = 0; r < n; r++) {
            const uint4 b0 = rows[r][0], b1 = rows[r][1];
            unsigned d = __popc(a0.x ^ b0.x);
            d += __popc(a0.y ^ b0.y);
            d += __popc(a0.z ^ b0.z);
            d += __popc(a0.w ^ b0.w);
            d += __popc(a1.x ^ b1.x);
            d += __popc(a1.y ^ b1.y);
            d += __popc(a1.z ^ b1.z);
            d += __popc(a1.w ^ b1.w);
            const unsigned key = (d << 20) | (unsigned)(t0 + r);
            k1 = med3u(k0, k1, key);
            k0 = min(k0, key);
        }
    }
    if (qi < nq) part[(size_t)qi * nchunks + chunk] = make_uint2(k0, k1);
}

// Bit -> int8 expansion of the MFMA operands.  Flat dword F (0..63) of a
// descriptor holds, in bytes 0..3, bits {i, i+8, i+16, i+24} of 32-bit word
// u = F/8, with i = F%8.  Map side: a set bit becomes byte 0x80 (-128), a
// clear bit 0 -- one shift and one mask.  Query side: a set bit becomes 0x40
// (+64), a clear bit 0xC0 (-64) -- the same with an XOR.
__device__ __forceinline__ unsigned expand_map_dword(unsigned w, int i)
{
    return (w << (7 - i)) & 0x80808080u;
}

__device__ __forceinline__ unsigned expand_query_dword(unsigned w, int i)
{
    return expand_map_dword(w, i) ^ 0xC0C0C0C0u;
}

typedef int knn_v4i __attribute__((ext_vector_type(4)));

// One 64-row tile into LDS: this thread's four map words (rows 16j + fl%16)
// expanded to the operand dwords F0..F0+3 of lane fl, k-chunk fc.
__device__ __forceinline__ void knn_fill_tile(uint4 (*A)[4][64], const unsigned* wn, int fc, int fl, int fi0)
{
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint4 v;
        v.x = expand_map_dword(wn[j], fi0 + 0);
        v.y = expand_map_dword(wn[j], fi0 + 1);
        v.z = expand_map_dword(wn[j], fi0 + 2);
        v.w = expand_map_dword(wn[j], fi0 + 3);
        A[j][fc][fl] = v;
    }
}

// The 16 MFMAs of subtile s (16 map rows x 64 queries): acc[gq] = C + sum
// over the four 64-wide k-chunks, C = `rows`.
__device__ __forceinline__ void knn_subtile(const uint4 (*A)[4][64], int s, int lane, const knn_v4i (*B)[4],
                                            knn_v4i rows, knn_v4i* acc)
{
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const uint4 av = A[s][c][lane];
        const knn_v4i a = (knn_v4i){(int)av.x, (int)av.y, (int)av.z, (int)av.w};
#pragma unroll
        for (int gq = 0; gq < 4; gq++)
            acc[gq] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, B[gq][c], c == 0 ? rows : acc[gq], 0, 0, 0);
    }
}

// Workgroup: 4 wavefronts x 64 queries, one map chunk.  Map rows are
// expanded 64 at a time into LDS in operand order [subtile][k-chunk][lane];
// each wavefront runs 4 (k-chunks) x 4 (query groups) MFMAs per 16 rows.
//
// The operands are scaled so that the accumulator is the search key itself.
// With map bits m -> -128 m and query bits q -> 64 (2q - 1),
//   sum -128 m 64 (2q - 1) = 8192 sum m (1 - 2q) = 8192 (Hamming - |q|)
// since Hamming = |q| + sum m (1 - 2q); the MFMA's C input is the row's index
// in the chunk.  |q| is constant per query, so
//   key = 8192 (Hamming - |q|) + row          (|key| < 2^22)
// orders (distance, row) exactly for rows under 8192 = KNN_CHUNK, and the
// top-2 update is one v_med3_i32 and one v_min_i32 per map row and query.
// 256 VGPRs would leave 3 wavefronts per SIMD and 768 of the 1024 workgroup
// slots for the 984 workgroups of a 2000 x 1M search; the bound of 4 keeps
// every workgroup resident.
__global__ __launch_bounds__(256, 4) void knn_mfma_kernel(const uint8_t* __restrict__ q, int nq,
                                                          const uint8_t* __restrict__ m, long long nm,
                                                          uint2* __restrict__ part, int nchunks)
{
    __shared__ uint4 A[4][4][64];   // 16 KB: 64 rows of expanded bytes
    __shared__ int popq[256];       // |q| of the workgroup's queries
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane >> 4, i16 = lane & 15;
    const int chunk = blockIdx.x;
    const int qbase = blockIdx.y * 256 + wv * 64;
    {
        const int qi = blockIdx.y * 256 + tid;
        int pc = 0;
        if (qi < nq) {
#pragma unroll
            for (int u = 0; u < 8; u++) pc += __popc(*(const unsigned*)(q + (size_t)qi * 32 + 4 * u));
        }
        popq[tid] = pc;   // read after the loop's first barrier
    }

    // Query operands: group gq, k-chunk c -> this lane's 4 dwords F = 16c+4g+d.
    knn_v4i B[4][4];
#pragma unroll
    for (int gq = 0; gq < 4; gq++) {
        const int qi = qbase + 16 * gq + i16;
        unsigned w[8];
#pragma unroll
        for (int u = 0; u < 8; u++) w[u] = qi < nq ? *(const unsigned*)(q + (size_t)qi * 32 + 4 * u) : 0u;
#pragma unroll
        for (int c = 0; c < 4; c++) {
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int F = 16 * c + 4 * g + d;
                B[gq][c][d] = (int)expand_query_dword(w[F >> 3], F & 7);
            }
        }
    }
    const long long c0 = (long long)chunk * KNN_CHUNK;
    const int len = (int)min((long long)KNN_CHUNK, nm - c0);
    int k0[4], k1[4];
#pragma unroll
    for (int gq = 0; gq < 4; gq++) k0[gq] = k1[gq] = INT_MAX;

    // Accumulator (lane, r) = (map row t0 + 16s + 4g + r, query 16gq + i16):
    // `rows` is the C input of subtile s, advanced by 16 per subtile.
    knn_v4i rows = (knn_v4i){4 * g, 4 * g + 1, 4 * g + 2, 4 * g + 3};
    // Fill role of this thread: k-chunk fc = wavefront, operand lane fl, so
    // that the wavefront's 16-byte LDS writes are consecutive (no bank
    // conflicts): for subtile j, row 16j + (fl & 15), dwords F0..F0+3 with
    // F0 = 16 fc + 4 (fl >> 4), i.e. bits 4 (fl >> 4 & 1) + 0..3 (+8k) of
    // word 2 fc + (fl >> 5).  The next tile's words are loaded before the
    // current tile's MFMAs run.  Full 64-row tiles run without bounds checks;
    // the chunk's last, partial tile takes the checked form.
    // Rows past the chunk's end read its last row instead (their keys are
    // discarded in the checked form).
    const int fc = wv, fl = lane;
    const int fi0 = 4 * ((fl >> 4) & 1);
    const uint8_t* mchunk = m + (size_t)c0 * 32;                 // uniform base
    const unsigned fword = 4u * (unsigned)(2 * fc + (fl >> 5));  // + 32-bit lane offsets
    const int nfull = len & ~63;
    unsigned wn[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
        wn[j] = *(const unsigned*)(mchunk + ((unsigned)min(16 * j + (fl & 15), len - 1) * 32u + fword));
    for (int t0 = 0; t0 < nfull; t0 += 64) {
        __syncthreads();
        knn_fill_tile(A, wn, fc, fl, fi0);
        __syncthreads();
        if (t0 + 64 < len) {
#pragma unroll
            for (int j = 0; j < 4; j++)
                wn[j] = *(const unsigned*)(mchunk + ((unsigned)min(t0 + 64 + 16 * j + (fl & 15), len - 1) * 32u + fword));
        }
#pragma unroll
        for (int s = 0; s < 4; s++) {
            knn_v4i acc[4];
            knn_subtile(A, s, lane, B, rows, acc);
            rows += 16;
#pragma unroll
            for (int r = 0; r < 4; r++) {
#pragma unroll
                for (int gq = 0; gq < 4; gq++) {
                    const int key = acc[gq][r];
                    k1[gq] = med3i(k0[gq], k1[gq], key);
                    k0[gq] = min(k0[gq], key);
                }
            }
        }
    }
    if (nfull < len) {
        __syncthreads();
        knn_fill_tile(A, wn, fc, fl, fi0);
        __syncthreads();
#pragma unroll 1
        for (int s = 0; 16 * s < len - nfull; s++) {
            knn_v4i acc[4];
            knn_subtile(A, s, lane, B, rows, acc);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const bool ok = rows[r] < len;
#pragma unroll
                for (int gq = 0; gq < 4; gq++) {
                    const int key = ok ? acc[gq][r] : INT_MAX;
                    k1[gq] = med3i(k0[gq], k1[gq], key);
                    k0[gq] = min(k0[gq], key);
                }
            }
            rows += 16;
        }
    }
    // Merge the four lanes that hold the same query (lanes i16 + 16g) and
    // convert to the partial format distance << 20 | row.
#pragma unroll
    for (int gq = 0; gq < 4; gq++) {
#pragma unroll
        for (int o = 16; o <= 32; o <<= 1) {
            const int b0 = __shfl_xor(k0[gq], o, 64), b1 = __shfl_xor(k1[gq], o, 64);
            const int n0 = min(k0[gq], b0);
            const int n1 = min(max(k0[gq], b0), min(k1[gq], b1));
            k0[gq] = n0;
            k1[gq] = n1;
        }
        const int qi = qbase + 16 * gq + i16;
        if (g == 0 && qi < nq) {
            const int pc = popq[wv * 64 + 16 * gq + i16];
            const int ks[2] = {k0[gq], k1[gq]};
            unsigned out[2];
#pragma unroll
            for (int t = 0; t < 2; t++)
                out[t] = ks[t] == INT_MAX ? KNN_NONE
                                          : ((unsigned)((ks[t] >> 13) + pc) << 20) | (unsigned)(ks[t] & (KNN_CHUNK - 1));
            part[(size_t)qi * nchunks + chunk] = make_uint2(out[0], out[1]);
        }
    }
}

// Reduce the per-chunk keys of each query to the global two nearest and apply
// the ratio test (best < ratio * second, in float).
__global__ __launch_bounds__(256) void knn_merge_kernel(const uint2* __restrict__ part, int nq, int nchunks,
                                                        float ratio, int* __restrict__ idx2, int* __restrict__ dist2,
                                                        uint8_t* __restrict__ accept)
{
    const int qi = blockIdx.x * 256 + threadIdx.x;
    if (qi >= nq) return;
    unsigned long long b0 = ~0ull, b1 = ~0ull;
    for (int c = 0; c < nchunks; c++) {
        const uint2 p = part[(size_t)qi * nchunks + c];
        const unsigned ks[2] = {p.x, p.y};
        for (int t = 0; t < 2; t++) {
            if (ks[t] == KNN_NONE) continue;
            const unsigned long long k = ((unsigned long long)(ks[t] >> 20) << 40) |
                                         (unsigned long long)((long long)c * KNN_CHUNK + (ks[t] & 0xFFFFF));
            if (k < b0) {
                b1 = b0;
                b0 = k;
            } else if (k < b1) {
                b1 = k;
            }
        }
    }
    const int d0 = b0 == ~0ull ? INT_MAX : (int)(b0 >> 40), d1 = b1 == ~0ull ? INT_MAX : (int)(b1 >> 40);
    idx2[2 * qi] = b0 == ~0ull ? -1 : (int)(b0 & 0xFFFFFFFFFFull);
    idx2[2 * qi + 1] = b1 == ~0ull ? -1 : (int)(b1 & 0xFFFFFFFFFFull);
    dist2[2 * qi] = d0;
    dist2[2 * qi + 1] = d1;
    if (accept) accept[qi] = (uint8_t)(b0 != ~0ull && b1 != ~0ull && (float)d0 < (float)d1 * ratio);
}

extern "C" int orb_knn_chunks(long long nm) { return (int)((nm + KNN_CHUNK - 1) / KNN_CHUNK); }

extern "C" hipError_t orb_launch_knn2(const uint8_t* q, int nq, const uint8_t* m, long long nm, int method,
                                      float ratio, uint2* part, int* idx2, int* dist2, uint8_t* accept,
                                      hipStream_t st, hipEvent_t* ev)
{
    const int nchunks = orb_knn_chunks(nm);
    if (nq <= 0) return hipSuccess;
    if (ev) (void)hipEventRecord(ev[0], st);
    if (nchunks > 0) {
        const dim3 grid(nchunks, (nq + 255) / 256);
        if (method == 1)
            knn_mfma_kernel<<<grid, 256, 0, st>>>(q, nq, m, nm, part, nchunks);
        else
            knn_popc_kernel<<<grid, 256, 0, st>>>(q, nq, m, nm, part, nchunks);
    }
    if (ev) (void)hipEventRecord(ev[1], st);
    knn_merge_kernel<<<(nq + 255) / 256, 256, 0, st>>>(part, nq, nchunks, ratio, idx2, dist2, accept);
    if (ev) (void)hipEventRecord(ev[2], st);
    return hipGetLastError();
}

// ================================================= allocation-free dense k=2
// orb_knn2_device_noalloc keeps the per-chunk partial keys in this fixed
// array instead of a context buffer, so that it makes no allocation.  Each
// GPU holds its own copy until the process ends; host code uses the address
// orb_knn_scratch_setup looked up for the device, never the symbol itself.
#define ORB_KNN_SCRATCH_ENTRIES (1 << 19)   // uint2 keys: 4 MiB per GPU
#define ORB_KNN_SCRATCH_DEVICES 64
static __device__ __align__(256) uint2 scratch[ORB_KNN_SCRATCH_ENTRIES];
static uint2* g_scratch_ptr[ORB_KNN_SCRATCH_DEVICES];

// With `device` current: look up this GPU's copy of the array (once).
// Returns 0, or a HIP error code.
extern "C" int orb_knn_scratch_setup(int device)
{
    if (device < 0 || device >= ORB_KNN_SCRATCH_DEVICES) return (int)hipErrorInvalidDevice;
    if (__atomic_load_n(&g_scratch_ptr[device], __ATOMIC_ACQUIRE)) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!ptr) return (int)hipErrorInvalidSymbol;
    __atomic_store_n(&g_scratch_ptr[device], (uint2*)ptr, __ATOMIC_RELEASE);
    return 0;
}

// The device address of `device`'s array (null before its setup) and the
// number of uint2 keys it holds.
extern "C" uint2* orb_knn_scratch(int device, size_t* entries)
{
    if (entries) *entries = ORB_KNN_SCRATCH_ENTRIES;
    if (device < 0 || device >= ORB_KNN_SCRATCH_DEVICES) return nullptr;
    return __atomic_load_n(&g_scratch_ptr[device], __ATOMIC_ACQUIRE);
}

extern "C" hipError_t orb_launch_match_pairs(const MatchArgs* d_args, int npairs, int nwaves, hipStream_t st)
{
    if (npairs <= 0) return hipSuccess;
    if (nwaves > 0) match_pairs_kernel<<<(nwaves + 3) / 4, 256, 0, st>>>(d_args, npairs, nwaves);
    save_prev_kernel<<<npairs, 256, 0, st>>>(d_args, npairs);
    return hipGetLastError();
}

// ================================================== batch staged transfers
// One launch for a list of regions (blockIdx.y = region): a batch's pinned
// input images, landed in one device staging buffer by one host-to-device
// copy, go to their slots' images; the slots' outputs go into one buffer
// for one device-to-host copy.  A region whose length is a count read on the
// device copies only the valid entries.
__global__ __launch_bounds__(256) void copy_ops_kernel(const CopyOp* __restrict__ ops)
{
    const CopyOp o = ops[blockIdx.y];
    int n = o.words;
    if (o.count) n = min(n, max(0, *o.count) * o.unit_words);
    const int stride = gridDim.x * 256;
    int i = blockIdx.x * 256 + threadIdx.x;
    if ((n & 3) == 0 && (((uintptr_t)o.src | (uintptr_t)o.dst) & 15) == 0) {
        const int n4 = n >> 2;
        for (; i < n4; i += stride) ((uint4*)o.dst)[i] = ((const uint4*)o.src)[i];
    } else {
        for (; i < n; i += stride) o.dst[i] = o.src[i];
    }
}

extern "C" hipError_t orb_launch_copy_ops(const CopyOp* d_ops, int nops, int blocks_per_op, hipStream_t st)
{
    if (nops <= 0) return hipSuccess;
    copy_ops_kernel<<<dim3(blocks_per_op, nops), 256, 0, st>>>(d_ops);
    return hipGetLastError();
}

// ======================================================= HBM copy microbench
// Plain streaming copy (16-byte loads and stores, grid-stride): the measured
// HBM rate of this kernel re-confirms the bandwidth peak, and its counter
// readings calibrate FETCH_SIZE / WRITE_SIZE for a known byte count.
__global__ __launch_bounds__(256) void hbm_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                       size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

extern "C" hipError_t orb_launch_hbm_copy(const void* src, void* dst, size_t bytes, hipStream_t st)
{
    const size_t n16 = bytes / 16;
    hbm_copy_kernel<<<8192, 256, 0, st>>>((const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError();
}
