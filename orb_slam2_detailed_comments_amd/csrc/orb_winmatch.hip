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
//   win_round_kernel   one round of the in-order pass resolved in parallel
//                      (below), 16 lanes per query
//   win_finish_kernel  per pair: vnMatches12, nmatches, the rotation check and
//                      vbPrevMatched from the settled round's outcomes
//   win_greedy_kernel  per pair (one wavefront): the reference's sequential
//                      pass over the queries, each query's list reduced
//                      across the lanes, state (vMatchedDistance,
//                      vnMatches21, vnMatches12, the rotation histogram) in
//                      LDS; for pairs whose rounds did not settle
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
    if (P.changed) {   // SearchForInitialization's rounds: rounds 0 and 1 write claim buffers 1 and 2
        for (int i = threadIdx.x; i < n; i += 1024) P.ccnt[0][i] = P.ccnt[1][i] = 0;
        for (int k = threadIdx.x; k < ORB_WIN_MAXIT; k += 1024) P.changed[k] = k == 0 ? 1 : 4;   // bit 2: not run
    }
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
// The first round k >= 1 that ran, changed no outcome and read complete
// claims (round k - 1 did not overflow its slots): its outcomes are the
// in-order pass's.  0 if there is none.
__device__ __forceinline__ int win_settled_round(const int* chg)
{
    for (int k = 1; k < ORB_WIN_MAXIT; k++)
        if (chg[k] == 0 && !(chg[k - 1] & 2)) return k;
    return 0;
}

__global__ __launch_bounds__(64) void win_greedy_kernel(const WinPair* __restrict__ pairs, int skip_settled)
{
    __shared__ int vmd[WIN_CAP];
    __shared__ int m21[WIN_CAP];
    __shared__ int m12[WIN_CAP];
    __shared__ signed char bin[WIN_CAP];
    __shared__ int hist[WIN_HISTO];
    __shared__ int s_n;
    const WinPair& P = pairs[blockIdx.x];
    if (skip_settled && win_settled_round(P.changed) > 0) return;
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

extern "C" hipError_t orb_launch_win_fill(const WinPair* d_pairs, int npairs, int max_n1, hipStream_t st)
{
    if (npairs <= 0 || max_n1 <= 0) return hipSuccess;
    win_fill_kernel<<<dim3((max_n1 + 3) / 4, npairs), 256, 0, st>>>(d_pairs);
    return hipGetLastError();
}

// ======================================================= SearchByProjection
// Both per-frame overloads of ORBmatcher::SearchByProjection
// (src/ORBmatcher.cc:72-183 with RadiusByViewingCos :186-202, and
// :1710-1889), many frames per call.
//
// The reference visits the points in order, and a keypoint claimed by an
// earlier point that has observations is skipped by later ones, so point p's
// outcome is a function of the outcomes of points < p.  Such a sequential
// pass is the unique fixed point of "recompute every outcome against the
// claims of the previous round's outcomes": by induction, round k fixes the
// first k points, and a round that changes nothing is that fixed point.
// Conflicts are local (a few px, one or two levels), so a handful of fully
// parallel rounds replace a pass over thousands of points.  A frame that has
// not settled after ORB_PROJ_MAXIT rounds gets the sequential pass.
//   proj_prep_*        the per-point arithmetic of each overload -> ProjQuery
//   win_grid_kernel    the frame's keypoint grid (as for SearchForInitialization)
//   proj_count_kernel  per query: candidates after the static filters (levels,
//                      circle, occupied keys, mvuRight)
//   win_scan_kernel    offsets
//   proj_fill_kernel   per query: (keypoint, distance, octave) in GetFeaturesInArea order
//   proj_round_kernel  one claim-resolution round, 16 lanes per query
//   proj_seq_kernel    the sequential pass, for frames that did not settle
//   proj_finish_kernel per frame: key_point, nmatches, the rotation check

// GetFeaturesInArea(x, y, r, minLevel, maxLevel) over the sorted grid; calls
// f(i2, position) in the reference's order for candidates that also pass
// keep(i2).  All lanes of the wavefront take part.
template <class K, class F>
__device__ __forceinline__ int area_enumerate(const ProjFrame& P, const ProjQuery& Q, K&& keep, F&& f)
{
    const int lane = threadIdx.x & 63;
    const float x = Q.x, y = Q.y, r = Q.r;
    const float invW = static_cast<float>(WIN_COLS) / static_cast<float>(P.maxX - P.minX);
    const float invH = static_cast<float>(WIN_ROWS) / static_cast<float>(P.maxY - P.minY);
    const int nMinCellX = max(0, (int)floorf((x - P.minX - r) * invW));
    if (nMinCellX >= WIN_COLS) return 0;
    const int nMaxCellX = min(WIN_COLS - 1, (int)ceilf((x - P.minX + r) * invW));
    if (nMaxCellX < 0) return 0;
    const int nMinCellY = max(0, (int)floorf((y - P.minY - r) * invH));
    if (nMinCellY >= WIN_ROWS) return 0;
    const int nMaxCellY = min(WIN_ROWS - 1, (int)ceilf((y - P.minY + r) * invH));
    if (nMaxCellY < 0) return 0;
    const bool checkLevels = (Q.minL > 0) || (Q.maxL >= 0);
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
                const OrbKeyPoint k = P.kp[i2];
                take = !checkLevels || (k.octave >= Q.minL && (Q.maxL < 0 || k.octave <= Q.maxL));
                if (take) {
                    const float distx = k.x - x, disty = k.y - y;
                    take = distx * distx + disty * disty < rr && keep(i2);
                }
            }
            const unsigned long long bal = __ballot(take);
            if (take) f(i2, pos + __popcll(bal & ((1ull << lane) - 1ull)));
            pos += __popcll(bal);
        }
    }
    return pos;
}

// Static candidate filters: keypoints held by a map point with observations,
// and mvuRight beyond the radius of the query's right-image u.
__device__ __forceinline__ bool proj_keep(const ProjFrame& P, const ProjQuery& Q, int i2)
{
    if (P.occupied[i2]) return false;
    const float ur = P.uright[i2];
    if (ur > 0) {
        const float er = fabsf(Q.su - ur);
        if (er > Q.r) return false;
    }
    return true;
}

// SearchByProjection(Frame&, vector<MapPoint*>, th) per point (:82-105):
// RadiusByViewingCos compares the float cosine with the double 0.998.
__global__ __launch_bounds__(256) void proj_prep_local_kernel(ProjFrame* __restrict__ frames,
                                                              const float* __restrict__ pts, const int* __restrict__ pt_off,
                                                              const float* __restrict__ scale, int nlevels, float th)
{
    ProjFrame& P = frames[blockIdx.y];
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < ORB_PROJ_MAXIT; k += 256) P.changed[k] = k == 0 ? 1 : 4;   // bit 2: not run
    if (q >= P.nq) return;
    const float* src = pts + (size_t)(pt_off[blockIdx.y] + q) * 6;   // proj_x, proj_y, proj_xr, view_cos, level, flags
    const int level = __float_as_int(src[4]), fl = __float_as_int(src[5]);
    ProjQuery Q;
    const bool valid = (fl & 1) && !(fl & 2) && level >= 0 && level < nlevels;
    float r = (double)src[3] > 0.998 ? 2.5f : 4.0f;
    if (th != 1.0f) r *= th;
    Q.x = src[0];
    Q.y = src[1];
    Q.r = valid ? r * scale[level] : 0.f;
    Q.su = src[2];
    Q.minL = level - 1;
    Q.maxL = level;
    Q.flags = (valid ? ORB_PQ_VALID : 0) | (fl & 4 ? ORB_PQ_OBS : 0);
    Q.angle = 0.f;
    P.q[q] = Q;
}

// SearchByProjection(Frame& Current, const Frame& Last, th, bMono) per point
// (:1743-1795): pinhole projection of the camera-frame point, bounds, radius
// and the level window chosen by the motion along the optical axis.
// cam per frame: fx, fy, cx, cy, mbf, mb, tlc_z.
__global__ __launch_bounds__(256) void proj_prep_last_kernel(ProjFrame* __restrict__ frames,
                                                             const float* __restrict__ pts, const int* __restrict__ pt_off,
                                                             const float* __restrict__ scale, int nlevels,
                                                             const float* __restrict__ cam, int mono, float th)
{
    ProjFrame& P = frames[blockIdx.y];
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < ORB_PROJ_MAXIT; k += 256) P.changed[k] = k == 0 ? 1 : 4;   // bit 2: not run
    if (q >= P.nq) return;
    const float* c = cam + (size_t)blockIdx.y * 7;
    const float fx = c[0], fy = c[1], cx = c[2], cy = c[3], mbf = c[4], mb = c[5], tlc_z = c[6];
    const bool bForward = tlc_z > mb && !mono;
    const bool bBackward = -tlc_z > mb && !mono;
    const float* src = pts + (size_t)(pt_off[blockIdx.y] + q) * 6;   // xc, yc, zc, angle, octave, flags
    const int octave = __float_as_int(src[4]), fl = __float_as_int(src[5]);
    ProjQuery Q;
    Q.flags = fl & 4 ? ORB_PQ_OBS : 0;
    Q.angle = src[3];
    Q.x = Q.y = Q.r = Q.su = 0.f;
    Q.minL = Q.maxL = -1;
    const float invzc = (float)(1.0 / (double)src[2]);
    bool valid = (fl & 1) && octave >= 0 && octave < nlevels && !(invzc < 0);
    if (valid) {
        const float u = fx * src[0] * invzc + cx;
        const float v = fy * src[1] * invzc + cy;
        valid = !(u < P.minX || u > P.maxX) && !(v < P.minY || v > P.maxY);
        Q.x = u;
        Q.y = v;
        Q.r = th * scale[octave];
        Q.su = u - mbf * invzc;
        if (bForward) {
            Q.minL = octave;
            Q.maxL = -1;
        } else if (bBackward) {
            Q.minL = 0;
            Q.maxL = octave;
        } else {
            Q.minL = octave - 1;
            Q.maxL = octave + 1;
        }
    }
    if (valid) Q.flags |= ORB_PQ_VALID;
    P.q[q] = Q;
}

// grid (ceil(max nq / 4), frames): one wavefront per query.  Also resets the
// first two claim buffers of the frame (no claims before round 0).
__global__ __launch_bounds__(256) void proj_count_kernel(ProjFrame* __restrict__ frames)
{
    ProjFrame& P = frames[blockIdx.y];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < P.n; i += gridDim.x * 256) {
        P.claim[0][i] = 0x7FFFFFFF;
        P.claim[1][i] = 0x7FFFFFFF;
    }
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= P.nq) return;
    const ProjQuery Q = P.q[q];
    int c = 0;
    if (Q.flags & ORB_PQ_VALID)
        c = area_enumerate(P, Q, [&](int i2) { return proj_keep(P, Q, i2); }, [](int, int) {});
    if ((threadIdx.x & 63) == 0) P.cnt[q] = c;
}

__global__ __launch_bounds__(256) void proj_fill_kernel(const ProjFrame* __restrict__ frames)
{
    const ProjFrame& P = frames[blockIdx.y];
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= P.nq || P.cnt[q] == 0) return;
    const ProjQuery Q = P.q[q];
    const uint4 a0 = *(const uint4*)(P.qdesc + (size_t)q * 32), a1 = *(const uint4*)(P.qdesc + (size_t)q * 32 + 16);
    unsigned* out = P.cand + P.off[q];
    area_enumerate(P, Q, [&](int i2) { return proj_keep(P, Q, i2); }, [&](int i2, int p) {
        const uint4 b0 = *(const uint4*)(P.desc + (size_t)i2 * 32), b1 = *(const uint4*)(P.desc + (size_t)i2 * 32 + 16);
        const int d = __popc(a0.x ^ b0.x) + __popc(a0.y ^ b0.y) + __popc(a0.z ^ b0.z) + __popc(a0.w ^ b0.w) +
                      __popc(a1.x ^ b1.x) + __popc(a1.y ^ b1.y) + __popc(a1.z ^ b1.z) + __popc(a1.w ^ b1.w);
        out[p] = (unsigned)i2 | ((unsigned)d << 12) | ((unsigned)P.kp[i2].octave << 21);
    });
}

// The per-point decision over its candidate list with blocked keypoints
// removed, by `width` lanes (lanes lane0 .. lane0 + width - 1 of the
// wavefront take part; all must call).  Keys are distance << 12 | list
// position, so the smallest is the first minimum in candidate order and the
// second smallest is the reference's bestDist2 holder (the first candidate
// other than the best with that distance).  Rule 1 (:120-170): accept iff
// best <= TH_HIGH and not (same level as the second and best > ratio *
// second).  Rule 2 (:1806-1834): accept iff best <= TH_HIGH.
template <int WIDTH, class B>
__device__ __forceinline__ int proj_decide(const ProjFrame& P, int q, int sub, B&& blocked)
{
    const int L = P.cnt[q];
    const unsigned* list = P.cand + P.off[q];
    unsigned k1 = 0xFFFFFFFFu, k2 = 0xFFFFFFFFu;
    for (int p = sub; p < L; p += WIDTH) {
        const unsigned c = list[p];
        if (blocked((int)(c & 0xFFFu))) continue;
        const unsigned key = (((c >> 12) & 0x1FFu) << 12) | (unsigned)p;
        if (key < k1) {
            k2 = k1;
            k1 = key;
        } else if (key < k2) {
            k2 = key;
        }
    }
#pragma unroll
    for (int o = WIDTH / 2; o > 0; o >>= 1) {
        const unsigned o1 = __shfl_xor(k1, o, WIDTH), o2 = __shfl_xor(k2, o, WIDTH);
        const unsigned lo = min(k1, o1), hi = max(k1, o1);
        k2 = min(hi, min(k2, o2));
        k1 = lo;
    }
    if (k1 == 0xFFFFFFFFu) return -1;
    const int best = (int)(k1 >> 12);
    if (best > 100) return -1;
    const unsigned c1 = list[k1 & 0xFFFu];
    if (P.rule == 1 && k2 != 0xFFFFFFFFu) {
        const unsigned c2 = list[k2 & 0xFFFu];
        const int best2 = (int)(k2 >> 12);
        if ((c1 >> 21) == (c2 >> 21) && (float)best > P.nnratio * (float)best2) return -1;
    }
    return (int)(c1 & 0xFFFu);
}

// Round k (grid (ceil(max nq / 16), frames), 16 lanes per query): outcomes
// against the claims of round k - 1, claims of these outcomes, and the reset
// of the buffer round k + 1 will fill.  A frame settled at an earlier round
// (one that ran and changed nothing) is left alone.  A round clears bit 2 of
// its flag (set: not run).
__global__ __launch_bounds__(256) void proj_round_kernel(ProjFrame* __restrict__ frames, int k)
{
    ProjFrame& P = frames[blockIdx.y];
    for (int j = 1; j < k; j++)
        if (P.changed[j] == 0) return;                 // settled at round j: later rounds are left alone
    if (k > 0 && blockIdx.x == 0 && threadIdx.x == 0) atomicAnd(&P.changed[k], ~4);
    const int* __restrict__ claimR = P.claim[k % 3];
    int* claimW = P.claim[(k + 1) % 3];
    int* claimC = P.claim[(k + 2) % 3];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < P.n; i += gridDim.x * 256) claimC[i] = 0x7FFFFFFF;
    const int q = blockIdx.x * 16 + (threadIdx.x >> 4), sub = threadIdx.x & 15;
    const bool active = q < P.nq;
    int res = -1;
    if (active && (P.q[q].flags & ORB_PQ_VALID) && P.cnt[q] > 0)
        res = proj_decide<16>(P, q, sub, [&](int i2) { return claimR[i2] < q; });
    if (active && sub == 0) {
        P.out[k & 1][q] = res;
        if (k > 0 && P.out[(k - 1) & 1][q] != res) atomicOr(&P.changed[k], 1);
        if (res >= 0 && (P.q[q].flags & ORB_PQ_OBS)) atomicMin(&claimW[res], q);
    }
}

// The reference's sequential pass (one wavefront per frame), for frames that
// have not settled: claims in LDS, recorded as soon as a point is accepted.
// Writes round 1's outcome buffer and marks round 1 settled.
__global__ __launch_bounds__(64) void proj_seq_kernel(ProjFrame* __restrict__ frames)
{
    __shared__ int claimer[WIN_CAP];
    ProjFrame& P = frames[blockIdx.x];
    bool settled = false;
    for (int k = 1; k < ORB_PROJ_MAXIT; k++) settled |= P.changed[k] == 0;
    if (settled) return;
    for (int i = threadIdx.x; i < P.n; i += 64) claimer[i] = 0;
    __syncthreads();
    for (int q = 0; q < P.nq; q++) {
        int res = -1;
        if ((P.q[q].flags & ORB_PQ_VALID) && P.cnt[q] > 0)
            res = proj_decide<64>(P, q, threadIdx.x, [&](int i2) { return claimer[i2] != 0; });
        if (threadIdx.x == 0) {
            P.out[1][q] = res;
            if (res >= 0 && (P.q[q].flags & ORB_PQ_OBS)) claimer[res] = 1;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) P.changed[1] = 0;
}

// ORBmatcher::ComputeThreeMaxima (:2035-2077) on bin sizes.
__device__ __forceinline__ void three_maxima(const int* hist, int* ind)
{
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
    ind[0] = ind1;
    ind[1] = ind2;
    ind[2] = ind3;
}

// One workgroup per frame: the settled round's outcomes become pt_match;
// key_point[i] = the last point that claimed keypoint i (-1 none); rule 2
// with the rotation check builds the 30-bin histogram of the claims and sets
// key_point to -2 for the claims outside its three fullest bins
// (:1838-1886), each of which takes one off nmatches.
__global__ __launch_bounds__(1024) void proj_finish_kernel(const ProjFrame* __restrict__ frames)
{
    __shared__ int owner[WIN_CAP];
    __shared__ int hist[WIN_HISTO];
    __shared__ int s_ind[3];
    __shared__ int s_count[2];
    const ProjFrame& P = frames[blockIdx.x];
    int kc = 1;
    while (kc < ORB_PROJ_MAXIT - 1 && P.changed[kc] != 0) kc++;
    const int* out = P.out[kc & 1];
    for (int i = threadIdx.x; i < P.n; i += 1024) owner[i] = -1;
    if (threadIdx.x < WIN_HISTO) hist[threadIdx.x] = 0;
    if (threadIdx.x < 2) s_count[threadIdx.x] = 0;
    __syncthreads();
    const bool rot = P.rule == 2 && P.check_orientation;
    const float factor = WIN_HISTO / 360.0f;
    int nacc = 0;
    for (int q = threadIdx.x; q < P.nq; q += 1024) {
        const int m = out[q];
        P.pt_match[q] = m;
        if (m < 0) continue;
        nacc++;
        atomicMax(&owner[m], q);
        if (rot) {
            float r = P.q[q].angle - P.kp[m].angle;
            if (r < 0.0f) r += 360.0f;
            int b = (int)roundf(r * factor);
            if (b == WIN_HISTO) b = 0;
            atomicAdd(&hist[b], 1);
        }
    }
    atomicAdd(&s_count[0], nacc);
    __syncthreads();
    if (rot) {
        if (threadIdx.x == 0) three_maxima(hist, s_ind);
        __syncthreads();
        int removed = 0;
        for (int q = threadIdx.x; q < P.nq; q += 1024) {
            const int m = out[q];
            if (m < 0) continue;
            float r = P.q[q].angle - P.kp[m].angle;
            if (r < 0.0f) r += 360.0f;
            int b = (int)roundf(r * factor);
            if (b == WIN_HISTO) b = 0;
            if (b == s_ind[0] || b == s_ind[1] || b == s_ind[2]) continue;
            owner[m] = -2;          // every writer stores -2: no ordering needed
            removed++;
        }
        atomicAdd(&s_count[1], removed);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < P.n; i += 1024) P.key_point[i] = owner[i];
    if (threadIdx.x == 0) *P.nmatch = s_count[0] - s_count[1];
}

extern "C" hipError_t orb_launch_proj_prep(ProjFrame* d_frames, int nframes, int max_nq, int rule, const float* d_pts,
                                           const int* d_pt_off, const float* d_scale, int nlevels, const float* d_cam,
                                           int mono, float th, hipStream_t st)
{
    if (nframes <= 0) return hipSuccess;
    const dim3 g((max(max_nq, 1) + 255) / 256, nframes);
    if (rule == 1) proj_prep_local_kernel<<<g, 256, 0, st>>>(d_frames, d_pts, d_pt_off, d_scale, nlevels, th);
    else proj_prep_last_kernel<<<g, 256, 0, st>>>(d_frames, d_pts, d_pt_off, d_scale, nlevels, d_cam, mono, th);
    return hipGetLastError();
}

// d_grid: one WinPair per frame carrying the grid fields (kp2, n2, bounds,
// sorted, cell_start) and the scan fields (n1 = queries, cnt, off).
extern "C" hipError_t orb_launch_proj_count(const WinPair* d_grid, ProjFrame* d_frames, int nframes, int max_nq,
                                            int* d_totals, hipStream_t st)
{
    if (nframes <= 0) return hipSuccess;
    win_grid_kernel<<<nframes, 1024, 0, st>>>(d_grid);
    proj_count_kernel<<<dim3((max(max_nq, 1) + 3) / 4, nframes), 256, 0, st>>>(d_frames);
    win_scan_kernel<<<nframes, 1024, 0, st>>>(d_grid, d_totals);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_proj_fill(ProjFrame* d_frames, int nframes, int max_nq, hipStream_t st)
{
    if (nframes <= 0 || max_nq <= 0) return hipSuccess;
    proj_fill_kernel<<<dim3((max_nq + 3) / 4, nframes), 256, 0, st>>>(d_frames);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_proj_rounds(ProjFrame* d_frames, int nframes, int max_nq, int k0, int k1,
                                             hipStream_t st)
{
    if (nframes <= 0) return hipSuccess;
    for (int k = k0; k < k1; k++)
        proj_round_kernel<<<dim3((max(max_nq, 1) + 15) / 16, nframes), 256, 0, st>>>(d_frames, k);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_proj_finish(ProjFrame* d_frames, int nframes, int sequential, hipStream_t st)
{
    if (nframes <= 0) return hipSuccess;
    if (sequential) proj_seq_kernel<<<nframes, 64, 0, st>>>(d_frames);
    proj_finish_kernel<<<nframes, 1024, 0, st>>>(d_frames);
    return hipGetLastError();
}

// ============================== SearchForInitialization in parallel rounds
// The reference's pass (ORBmatcher.cc:599-665) visits the F1 keypoints in
// order; query i1 skips a candidate i2 whose vMatchedDistance is <= the
// candidate's distance, and an accepted query takes i2 (stealing it from its
// holder).  A target's successive holders have strictly decreasing
// distances, so vMatchedDistance[i2] before query i1 is the smallest distance
// of the queries < i1 accepted with target i2.  Query i1's outcome is thus a
// function of the outcomes of the queries before it, and the in-order pass
// is the unique fixed point of "recompute every outcome against the claims
// of the previous round's outcomes" (round k fixes the first k queries).
// Each target keeps up to ORB_WIN_SLOTS claims (query, distance) per round;
// a round that overflows them is not trusted as the fixed point.

// Round k (grid (ceil(max n1 / 16), pairs), 16 lanes per query): outcomes
// against the claims of round k - 1, claims of these outcomes, and the reset
// of the counts round k + 1 will fill.  A pair settled at an earlier round is
// left alone.  A round clears bit 2 of its flag (set: not run).
__global__ __launch_bounds__(256) void win_round_kernel(const WinPair* __restrict__ pairs, int k)
{
    const WinPair& P = pairs[blockIdx.y];
    const int ks = win_settled_round(P.changed);
    if (ks > 0 && ks < k) return;
    if (k > 0 && blockIdx.x == 0 && threadIdx.x == 0) atomicAnd(&P.changed[k], ~4);
    const int* __restrict__ cntR = P.ccnt[k % 3];
    const unsigned* __restrict__ claimR = P.claim[k % 3];
    int* cntW = P.ccnt[(k + 1) % 3];
    unsigned* claimW = P.claim[(k + 1) % 3];
    int* cntC = P.ccnt[(k + 2) % 3];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < P.n2; i += gridDim.x * 256) cntC[i] = 0;
    const int q = blockIdx.x * 16 + (threadIdx.x >> 4), sub = threadIdx.x & 15;
    const bool active = q < P.n1;
    int res = -1;
    if (active && P.cnt[q] > 0 && P.kp1[q].octave == 0) {   // the 16 lanes of a query agree
        const int L = P.cnt[q];
        const int2* list = P.cand + P.off[q];
        unsigned b1 = 0xFFFFFFFFu;
        int b2 = 0x7FFFFFFF;
        for (int p = sub; p < L; p += 16) {
            const int2 c = list[p];
            const int nc = min(cntR[c.x], ORB_WIN_SLOTS);
            bool blocked = false;
            for (int j = 0; j < nc; j++) {
                const unsigned e = claimR[(size_t)c.x * ORB_WIN_SLOTS + j];
                blocked |= (int)(e >> 9) < q && (int)(e & 511u) <= c.y;
            }
            if (blocked) continue;
            const unsigned key = ((unsigned)c.y << 12) | (unsigned)p;
            if (key < b1) {
                if (b1 != 0xFFFFFFFFu) b2 = min(b2, (int)(b1 >> 12));
                b1 = key;
            } else if (c.y < b2) {
                b2 = c.y;
            }
        }
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            const unsigned ob1 = __shfl_xor(b1, o, 16);
            const int ob2 = __shfl_xor(b2, o, 16);
            const unsigned hi = max(b1, ob1);
            int nb2 = min(b2, ob2);
            if (hi != 0xFFFFFFFFu) nb2 = min(nb2, (int)(hi >> 12));
            b1 = min(b1, ob1);
            b2 = nb2;
        }
        if (b1 != 0xFFFFFFFFu) {
            const int bestDist = (int)(b1 >> 12);
            if (bestDist <= 50 && (float)bestDist < (float)b2 * P.nnratio) res = (list[b1 & 0xFFFu].x << 9) | bestDist;
        }
    }
    if (active && sub == 0) {
        P.out[k & 1][q] = res;
        if (k > 0 && P.out[(k - 1) & 1][q] != res) atomicOr(&P.changed[k], 1);
        if (res >= 0) {
            const int t = res >> 9;
            const int slot = atomicAdd(&cntW[t], 1);
            if (slot < ORB_WIN_SLOTS) claimW[(size_t)t * ORB_WIN_SLOTS + slot] = ((unsigned)q << 9) | (unsigned)(res & 511);
            else atomicOr(&P.changed[k], 2);
        }
    }
}

// One workgroup per settled pair: every accepted query is a match event (a
// rotHist entry); a target stays with its last holder (vnMatches12 of the
// others was reset by the steal); the rotation check drops the matches whose
// event lies outside the three fullest bins (:667-690); vbPrevMatched takes
// the kept matches' F2 positions (:693-695).
__global__ __launch_bounds__(1024) void win_finish_kernel(const WinPair* __restrict__ pairs)
{
    __shared__ int owner[WIN_CAP];
    __shared__ signed char bin[WIN_CAP];
    __shared__ int hist[WIN_HISTO];
    __shared__ int s_ind[3];
    __shared__ int s_n[2];
    const WinPair& P = pairs[blockIdx.x];
    const int ks = win_settled_round(P.changed);
    if (ks == 0) return;                           // the sequential pass takes this pair
    const int* out = P.out[ks & 1];
    const int n1 = P.n1, t = threadIdx.x;
    for (int i = t; i < P.n2; i += 1024) owner[i] = -1;
    if (t < WIN_HISTO) hist[t] = 0;
    if (t < 2) s_n[t] = 0;
    __syncthreads();
    for (int q = t; q < n1; q += 1024) {
        const int r = out[q];
        if (r >= 0) atomicMax(&owner[r >> 9], q);
    }
    __syncthreads();
    const float factor = WIN_HISTO / 360.0f;
    int kept = 0;
    for (int q = t; q < n1; q += 1024) {
        const int r = out[q];
        int m = -1, b = -1;
        if (r >= 0) {
            const int tg = r >> 9;
            if (owner[tg] == q) {
                m = tg;
                kept++;
            }
            if (P.check_orientation) {
                float rot = P.kp1[q].angle - P.kp2[tg].angle;
                if (rot < 0.0f) rot += 360.0f;
                b = (int)roundf(rot * factor);
                if (b == WIN_HISTO) b = 0;
                atomicAdd(&hist[b], 1);
            }
        }
        bin[q] = (signed char)b;
        P.m12[q] = m;
    }
    atomicAdd(&s_n[0], kept);
    __syncthreads();
    if (P.check_orientation) {
        if (t == 0) three_maxima(hist, s_ind);
        __syncthreads();
        int dropped = 0;
        for (int q = t; q < n1; q += 1024) {
            const int b = bin[q];
            if (b < 0 || b == s_ind[0] || b == s_ind[1] || b == s_ind[2]) continue;
            if (P.m12[q] >= 0) {
                P.m12[q] = -1;
                dropped++;
            }
        }
        atomicAdd(&s_n[1], dropped);
        __syncthreads();
    }
    for (int q = t; q < n1; q += 1024) {
        const int m = P.m12[q];
        if (m >= 0) P.prev[q] = make_float2(P.kp2[m].x, P.kp2[m].y);
    }
    if (t == 0) *P.nmatch = s_n[0] - s_n[1];
}

extern "C" hipError_t orb_launch_win_rounds(const WinPair* d_pairs, int npairs, int max_n1, int k0, int k1,
                                            hipStream_t st)
{
    if (npairs <= 0) return hipSuccess;
    for (int k = k0; k < k1; k++)
        win_round_kernel<<<dim3((max(max_n1, 1) + 15) / 16, npairs), 256, 0, st>>>(d_pairs, k);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_win_finish(const WinPair* d_pairs, int npairs, int sequential, hipStream_t st)
{
    if (npairs <= 0) return hipSuccess;
    win_finish_kernel<<<npairs, 1024, 0, st>>>(d_pairs);
    if (sequential) win_greedy_kernel<<<npairs, 64, 0, st>>>(d_pairs, 1);
    return hipGetLastError();
}
