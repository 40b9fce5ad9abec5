// orb_mappoint.hip -- MapPoint::ComputeDistinctiveDescriptors
// (src/MapPoint.cc:365-450) for many map points in one launch.
//
// One wavefront per point.  The reference fills the N x N distance matrix,
// sorts every row (the zero self-distance included) and takes element
// (size_t)(0.5*(N-1)); the first row with the smallest such median wins.
// Here, for row i, lane j holds d(i, j) for j = lane + 64c (c < 16, so up to
// 1024 observations), and the median is the smallest v with
// #{j : d(i, j) <= v} >= (N-1)/2 + 1, found by binary search over v in
// [0, 256] with ballot counts -- the same order statistic without a sort.
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MP_MAX_CHUNKS 16             // 16 x 64 = 1024 observations per point

__global__ __launch_bounds__(256) void distinctive_kernel(const uint8_t* __restrict__ desc,
                                                          const int* __restrict__ offsets, int n_points,
                                                          int* __restrict__ best, uint8_t* __restrict__ out_desc)
{
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= n_points) return;
    const int a = offsets[p], N = offsets[p + 1] - a;
    if (N <= 0) {                              // no observation: -1 and a zero row
        if (lane == 0) best[p] = -1;
        if (out_desc && lane < 2) ((uint4*)(out_desc + (size_t)p * 32))[lane] = make_uint4(0, 0, 0, 0);
        return;
    }
    const int nc = (N + 63) >> 6;
    const int need = (N - 1) / 2 + 1;          // rank of the median, 1-based
    const uint4* rows = (const uint4*)(desc + (size_t)a * 32);
    int bestMedian = 0x7FFFFFFF, bestIdx = 0;
    for (int i = 0; i < N; i++) {
        const uint4 r0 = rows[2 * i], r1 = rows[2 * i + 1];
        int d[MP_MAX_CHUNKS];
#pragma unroll
        for (int c = 0; c < MP_MAX_CHUNKS; c++) {
            const int j = lane + 64 * c;
            d[c] = 1024;                       // beyond every threshold: not counted
            if (c < nc && j < N) {
                const uint4 q0 = rows[2 * j], q1 = rows[2 * j + 1];
                d[c] = __popc(r0.x ^ q0.x) + __popc(r0.y ^ q0.y) + __popc(r0.z ^ q0.z) + __popc(r0.w ^ q0.w) +
                       __popc(r1.x ^ q1.x) + __popc(r1.y ^ q1.y) + __popc(r1.z ^ q1.z) + __popc(r1.w ^ q1.w);
            }
        }
        int lo = 0, hi = 256;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            int cnt = 0;
#pragma unroll
            for (int c = 0; c < MP_MAX_CHUNKS; c++)
                if (c < nc) cnt += __popcll(__ballot(d[c] <= mid));
            if (cnt >= need) hi = mid;
            else lo = mid + 1;
        }
        if (lo < bestMedian) {
            bestMedian = lo;
            bestIdx = i;
        }
    }
    if (lane == 0) best[p] = bestIdx;
    if (out_desc && lane < 2) ((uint4*)(out_desc + (size_t)p * 32))[lane] = rows[2 * bestIdx + lane];
}

extern "C" hipError_t orb_launch_distinctive(const uint8_t* desc, const int* offsets, int n_points, int* best,
                                             uint8_t* out_desc, hipStream_t st)
{
    if (n_points <= 0) return hipSuccess;
    distinctive_kernel<<<(n_points + 3) / 4, 256, 0, st>>>(desc, offsets, n_points, best, out_desc);
    return hipGetLastError();
}
