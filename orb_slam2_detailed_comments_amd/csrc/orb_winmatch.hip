// orb_winmatch.hip -- ORBmatcher::SearchForInitialization
// (src/ORBmatcher.cc:573-717) over the Frame keypoint grid
// (Frame::AssignFeaturesToGrid / PosInGrid / GetFeaturesInArea,
// src/Frame.cc:399-423, 590-700), for many frame pairs per call.
//
// The reference visits F1's keypoints in order, and a candidate i2 is skipped
// when an earlier match already holds it at a distance <= this one
// (vMatchedDistance), so each query depends on all earlier ones.  The work
// splits into
//   win_grid_kernel    F2's grid: keypoints sorted by (cell = ix*48 + iy,
//                      index), so the cells of one grid column ix are one
//                      contiguous run, and GetFeaturesInArea's order (ix
//                      outer, iy inner, index order within a cell) is the
//                      sorted order over the window's columns
//   win_count_kernel   per query (one wavefront): how many candidates
//   win_scan_kernel    per pair: exclusive scan of the counts
//   win_fill_kernel    per query: the ordered (i2, distance) candidate list
//   win_greedy_kernel  per pair (one wavefront): the reference's sequential
//                      pass over the queries, each query's list reduced
//                      across the lanes, state (vMatchedDistance,
//                      vnMatches21, vnMatches12, the rotation histogram) in LDS
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "orb_types.h"

#define WIN_COLS 64
#define WIN_ROWS 48
#define WIN_CELLS (WIN_COLS * WIN_ROWS)
#define WIN_CAP 4096                 // keypoints per frame
#define WIN_HISTO 30

// Candidate enumeration of GetFeaturesInArea(x, y, r, 0, 0) over the sorted
// grid; calls f(i2, position) for each candidate in the reference's order.
// All lanes of the wavefront take part; positions are assigned by ballot.
template <class F>
__device__ __forceinline__ int win_enumerate(const WinPair& P, float x, float y, float r, F&& f)
{
    const int lane = threadIdx.x & 63;
    const float minX = P.minX, minY = P.minY;
    const float invW = static_cast<float>(WIN_COLS) / static_cast<float>(P.maxX - P.minX);
    const float invH = static_cast<float>(WIN_ROWS) / static_cast<float>(P.maxY - P.minY);
    const int nMinCellX = max(0, (int)floorf((x - minX - r) * invW));
    if (nMinCellX >= WIN_COLS) return 0;
    const int nMaxCellX = min(WIN_COLS - 1, (int)ceilf((x - minX + r) * invW));
    if (nMaxCellX < 0) return 0;
    const int nMinCellY = max(0, (int)floorf((y - minY - r) * invH));
    if (nMinCellY >= WIN_ROWS) return 0;
    const int nMaxCellY = min(WIN_ROWS - 1, (int)ceilf((y - minY + r) * invH));
    if (nMaxCellY < 0) return 0;
    const float rr = r * r;
    int pos = 0;
    for (int ix = nMinCellX; ix <= nMaxCellX; ix++) {
        const int a = P.cell_start[ix * WIN_ROWS + nMinCellY], b = P.cell_start[ix * WIN_ROWS + nMaxCellY + 1];
        for (int j0 = a; j0 < b; j0 += 64) {
            const int j = j0 + lane;
            bool take = false;
            int i2 = 0;
            if (j < b) {
                i2 = P.sorted[j];
                const OrbKeyPoint k = P.kp2[i2];
                if (k.octave == 0) {             // minLevel = maxLevel = level1 = 0
                    const float distx = k.x - x, disty = k.y - y;
                    take = distx * distx + disty * disty < rr;
                }
            }
            const unsigned long long bal = __ballot(take);
            if (take) {
                const int p = pos + __popcll(bal & ((1ull << lane) - 1ull));
                f(i2, p);
            }
            pos += __popcll(bal);
        }
    }
    return pos;
}

// One workgroup per pair: cell of every F2 keypoint (PosInGrid: round() of
// the float offset times the float inverse cell size; outside -> no cell),
// bitonic sort of (cell << 20 | index), and the first sorted position of
// every cell.
__global__ __launch_bounds__(1024) void win_grid_kernel(const WinPair* __restrict__ pairs)
{
    __shared__ unsigned key[WIN_CAP];
    const WinPair& P = pairs[blockIdx.x];
    const int n = P.n2;
    const float invW = static_cast<float>(WIN_COLS) / static_cast<float>(P.maxX - P.minX);
    const float invH = static_cast<float>(WIN_ROWS) / static_cast<float>(P.maxY - P.minY);
    int npow = 1;
    while (npow < n) npow <<= 1;
    for (int i = threadIdx.x; i < npow; i += 1024) {
        unsigned k = 0xFFFFFFFFu;
        if (i < n) {
            const OrbKeyPoint kp = P.kp2[i];
            const int px = (int)roundf((kp.x - P.minX) * invW), py = (int)roundf((kp.y - P.minY) * invH);
            const bool in = px >= 0 && px < WIN_COLS && py >= 0 && py < WIN_ROWS;
            k = ((unsigned)(in ? px * WIN_ROWS + py : 4095) << 20) | (unsigned)i;
        }
        key[i] = k;
    }
    __syncthreads();
    for (int k = 2; k <= npow; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int a = threadIdx.x; a < npow; a += 1024) {
                const int b = a ^ j;
                if (b > a) {
                    const unsigned va = key[a], vb = key[b];
                    const bool up = (a & k) == 0;
                    if (up ? (va > vb) : (va < vb)) {
                        key[a] = vb;
                        key[b] = va;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < n; i += 1024) P.sorted[i] = (int)(key[i] & 0xFFFFFu);
    for (int c = threadIdx.x; c <= WIN_CELLS; c += 1024) {
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)(key[mid] >> 20) < c) lo = mid + 1;
            else hi = mid;
        }
        P.cell_start[c] = lo;
    }
}

__device__ __forceinline__ bool win_query(const WinPair& P, int i1, float* x, float* y)
{
    if (P.kp1[i1].octave > 0) return false;
    *x = P.prev[i1].x;
    *y = P.prev[i1].y;
    return true;
}

// grid (ceil(max n1 / 4), pairs): one wavefront per F1 keypoint
__global__ __launch_bounds__(256) void win_count_kernel(const WinPair* __restrict__ pairs)
{
    const WinPair& P = pairs[blockIdx.y];
    const int i1 = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i1 >= P.n1) return;
    float x, y;
    int c = 0;
    if (win_query(P, i1, &x, &y)) c = win_enumerate(P, x, y, (float)P.window, [](int, int) {});
    if ((threadIdx.x & 63) == 0) P.cnt[i1] = c;
}

// one workgroup per pair: off[i] = sum of cnt[0..i), total in off[n1]
__global__ __launch_bounds__(1024) void win_scan_kernel(const WinPair* __restrict__ pairs, int* __restrict__ totals)
{
    __shared__ int part[1024];
    const WinPair& P = pairs[blockIdx.x];
    const int n = P.n1, t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int a = min(t * per, n), b = min(a + per, n);
    int s = 0;
    for (int i = a; i < b; i++) s += P.cnt[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = a; i < b; i++) {
        P.off[i] = run;
        run += P.cnt[i];
    }
    if (t == 1023) {
        P.off[n] = part[1023];
        totals[blockIdx.x] = part[1023];
    }
}

__global__ __launch_bounds__(256) void win_fill_kernel(const WinPair* __restrict__ pairs)
{
    const WinPair& P = pairs[blockIdx.y];
    const int i1 = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i1 >= P.n1) return;
    float x, y;
    if (!win_query(P, i1, &x, &y) || P.cnt[i1] == 0) return;
    const uint4 a0 = *(const uint4*)(P.d1 + (size_t)i1 * 32), a1 = *(const uint4*)(P.d1 + (size_t)i1 * 32 + 16);
    int2* out = P.cand + P.off[i1];
    win_enumerate(P, x, y, (float)P.window, [&](int i2, int p) {
        const uint4 b0 = *(const uint4*)(P.d2 + (size_t)i2 * 32), b1 = *(const uint4*)(P.d2 + (size_t)i2 * 32 + 16);
        const int d = __popc(a0.x ^ b0.x) + __popc(a0.y ^ b0.y) + __popc(a0.z ^ b0.z) + __popc(a0.w ^ b0.w) +
                      __popc(a1.x ^ b1.x) + __popc(a1.y ^ b1.y) + __popc(a1.z ^ b1.z) + __popc(a1.w ^ b1.w);
        out[p] = make_int2(i2, d);
    });
}

// One wavefront per pair.  Query by query, in index order: lanes take the
// list positions lane, lane+64, ... (ascending), skip candidates already
// held at a distance <= theirs, and keep (first minimum as distance << 12 |
// position, second smallest distance); a butterfly merge gives the
// reference's sequential best / second.  Lane 0 applies TH_LOW, the ratio
// and the steal of a matched target; the rotation histogram counts every
// match event, as rotHist's sizes do.
__global__ __launch_bounds__(64) void win_greedy_kernel(const WinPair* __restrict__ pairs)
{
    __shared__ int vmd[WIN_CAP];
    __shared__ int m21[WIN_CAP];
    __shared__ int m12[WIN_CAP];
    __shared__ signed char bin[WIN_CAP];
    __shared__ int hist[WIN_HISTO];
    __shared__ int s_n;
    const WinPair& P = pairs[blockIdx.x];
    const int lane = threadIdx.x;
    const int n1 = P.n1, n2 = P.n2;
    for (int i = lane; i < n2; i += 64) {
        vmd[i] = 0x7FFFFFFF;
        m21[i] = -1;
    }
    for (int i = lane; i < n1; i += 64) {
        m12[i] = -1;
        bin[i] = -1;
    }
    if (lane < WIN_HISTO) hist[lane] = 0;
    if (lane == 0) s_n = 0;
    __syncthreads();
    const float factor = WIN_HISTO / 360.0f;
    for (int i1 = 0; i1 < n1; i1++) {
        const int L = P.cnt[i1];
        if (L == 0 || P.kp1[i1].octave > 0) continue;
        const int2* list = P.cand + P.off[i1];
        unsigned b1 = 0xFFFFFFFFu;
        int b2 = 0x7FFFFFFF;
        for (int p = lane; p < L; p += 64) {
            const int2 c = list[p];
            if (vmd[c.x] <= c.y) continue;
            const unsigned key = ((unsigned)c.y << 12) | (unsigned)p;
            if (key < b1) {
                if (b1 != 0xFFFFFFFFu) b2 = min(b2, (int)(b1 >> 12));
                b1 = key;
            } else if (c.y < b2) {
                b2 = c.y;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned ob1 = __shfl_xor(b1, o, 64);
            const int ob2 = __shfl_xor(b2, o, 64);
            const unsigned hi = max(b1, ob1);
            int nb2 = min(b2, ob2);
            if (hi != 0xFFFFFFFFu) nb2 = min(nb2, (int)(hi >> 12));
            b1 = min(b1, ob1);
            b2 = nb2;
        }
        if (b1 == 0xFFFFFFFFu) continue;
        const int bestDist = (int)(b1 >> 12);
        if (lane == 0 && bestDist <= 50 && (float)bestDist < (float)b2 * P.nnratio) {
            const int bestIdx2 = list[b1 & 0xFFFu].x;
            if (m21[bestIdx2] >= 0) {
                m12[m21[bestIdx2]] = -1;
                s_n--;
            }
            m12[i1] = bestIdx2;
            m21[bestIdx2] = i1;
            vmd[bestIdx2] = bestDist;
            s_n++;
            if (P.check_orientation) {
                float rot = P.kp1[i1].angle - P.kp2[bestIdx2].angle;
                if (rot < 0.0f) rot += 360.0f;
                int b = (int)roundf(rot * factor);
                if (b == WIN_HISTO) b = 0;
                bin[i1] = (signed char)b;
                hist[b]++;
            }
        }
        __syncthreads();
    }
    if (P.check_orientation) {
        // ORBmatcher::ComputeThreeMaxima (:2035-2077) on the bin sizes
        __shared__ int s_ind[3];
        if (lane == 0) {
            int max1 = 0, max2 = 0, max3 = 0, ind1 = -1, ind2 = -1, ind3 = -1;
            for (int i = 0; i < WIN_HISTO; i++) {
                const int s = hist[i];
                if (s > max1) {
                    max3 = max2;
                    max2 = max1;
                    max1 = s;
                    ind3 = ind2;
                    ind2 = ind1;
                    ind1 = i;
                } else if (s > max2) {
                    max3 = max2;
                    max2 = s;
                    ind3 = ind2;
                    ind2 = i;
                } else if (s > max3) {
                    max3 = s;
                    ind3 = i;
                }
            }
            if (max2 < 0.1f * (float)max1) {
                ind2 = -1;
                ind3 = -1;
            } else if (max3 < 0.1f * (float)max1) {
                ind3 = -1;
            }
            s_ind[0] = ind1;
            s_ind[1] = ind2;
            s_ind[2] = ind3;
        }
        __syncthreads();
        int dropped = 0;
        for (int i = lane; i < n1; i += 64) {
            const int b = bin[i];
            if (b < 0 || b == s_ind[0] || b == s_ind[1] || b == s_ind[2]) continue;
            if (m12[i] >= 0) {
                m12[i] = -1;
                dropped++;
            }
        }
        for (int o = 32; o > 0; o >>= 1) dropped += __shfl_xor(dropped, o, 64);
        if (lane == 0) s_n -= dropped;
        __syncthreads();
    }
    for (int i = lane; i < n1; i += 64) {
        const int m = m12[i];
        P.m12[i] = m;
        if (m >= 0) P.prev[i] = make_float2(P.kp2[m].x, P.kp2[m].y);
    }
    if (lane == 0) *P.nmatch = s_n;
}

extern "C" hipError_t orb_launch_win_count(const WinPair* d_pairs, int npairs, int max_n1, hipStream_t st)
{
    if (npairs <= 0) return hipSuccess;
    win_grid_kernel<<<npairs, 1024, 0, st>>>(d_pairs);
    if (max_n1 > 0) win_count_kernel<<<dim3((max_n1 + 3) / 4, npairs), 256, 0, st>>>(d_pairs);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_win_scan(const WinPair* d_pairs, int npairs, int* d_totals, hipStream_t st)
{
    if (npairs <= 0) return hipSuccess;
    win_scan_kernel<<<npairs, 1024, 0, st>>>(d_pairs, d_totals);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_win_match(const WinPair* d_pairs, int npairs, int max_n1, hipStream_t st)
{
    if (npairs <= 0) return hipSuccess;
    if (max_n1 > 0) win_fill_kernel<<<dim3((max_n1 + 3) / 4, npairs), 256, 0, st>>>(d_pairs);
    win_greedy_kernel<<<npairs, 64, 0, st>>>(d_pairs);
    return hipGetLastError();
}
