// orb_frame.hip -- the per-keypoint Frame ops of the monocular and RGB-D
// constructors: Frame::UndistortKeyPoints (src/Frame.cc:724-776) and
// Frame::ComputeStereoFromRGBD (src/Frame.cc:1085-1123).  One thread per
// keypoint.  Bound: latency (a few thousand keypoints per frame).
#include <hip/hip_runtime.h>

#include "orb_types.h"

// cv::undistortPoints(pts, pts, K, D, Mat(), K) as OpenCV 2.4/3.x states it
// (cvUndistortPoints): in double, x = (u - cx) * (1/fx), five fixed-point
// iterations x = (x0 - deltaX) * icdist, then u = fx * x + cx (P = K, R = I),
// cast to float.  Built with contraction off, so every product and sum rounds
// as written.  k = k1 k2 p1 p2 k3 (k3 = 0 for a 4-coefficient model).
__global__ __launch_bounds__(256) void undistort_kernel(const OrbKeyPoint* __restrict__ in, int n,
                                                        const int* __restrict__ count, double fx,
                                                        double fy, double cx, double cy, double k0, double k1,
                                                        double k2, double k3, double k4, OrbKeyPoint* __restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= (count ? min(n, *count) : n)) return;
    OrbKeyPoint kp = in[i];
    const double ifx = 1. / fx, ify = 1. / fy;
    double x = ((double)kp.x - cx) * ifx;
    double y = ((double)kp.y - cy) * ify;
    const double x0 = x, y0 = y;
    for (int j = 0; j < 5; j++) {
        const double r2 = x * x + y * y;
        const double icdist = 1. / (1 + ((k4 * r2 + k1) * r2 + k0) * r2);
        const double deltaX = 2 * k2 * x * y + k3 * (r2 + 2 * x * x);
        const double deltaY = k2 * (r2 + 2 * y * y) + 2 * k3 * x * y;
        x = (x0 - deltaX) * icdist;
        y = (y0 - deltaY) * icdist;
    }
    kp.x = (float)(fx * x + cx);
    kp.y = (float)(fy * y + cy);
    out[i] = kp;
}

// d = depth[(int)y][(int)x] of the keypoint as extracted (Mat::at<float>
// truncates its float arguments); d > 0 gives depth d and uRight = xun - bf/d
// (float), else both -1.  Keypoints outside the image get -1.
__global__ __launch_bounds__(256) void rgbd_depth_kernel(const OrbKeyPoint* __restrict__ kps,
                                                         const OrbKeyPoint* __restrict__ kun, int n,
                                                         const int* __restrict__ count,
                                                         const float* __restrict__ depth, int w, int h,
                                                         size_t pitch, float bf, float* __restrict__ uR,
                                                         float* __restrict__ dep)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= (count ? min(n, *count) : n)) return;
    float r_out = -1.f, d_out = -1.f;
    const float u = kps[i].x, v = kps[i].y;
    if (v >= 0.f && u >= 0.f) {
        const int r = (int)v, c = (int)u;
        if (r < h && c < w) {
            const float d = *(const float*)((const char*)depth + (size_t)r * pitch + (size_t)c * 4);
            if (d > 0.f) {
                d_out = d;
                r_out = kun[i].x - __fdiv_rn(bf, d);
            }
        }
    }
    uR[i] = r_out;
    dep[i] = d_out;
}

// n: the keypoints, or (count set) the capacity of a list whose length the
// device holds in *count.
extern "C" hipError_t orb_launch_undistort(const OrbKeyPoint* in, int n, const int* count, const double* K4,
                                           const double* k5, OrbKeyPoint* out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    undistort_kernel<<<(n + 255) / 256, 256, 0, st>>>(in, n, count, K4[0], K4[1], K4[2], K4[3], k5[0], k5[1], k5[2],
                                                      k5[3], k5[4], out);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_rgbd_depth(const OrbKeyPoint* kps, const OrbKeyPoint* kun, int n, const int* count,
                                            const float* depth, int w, int h, size_t pitch, float bf, float* uR,
                                            float* dep, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    rgbd_depth_kernel<<<(n + 255) / 256, 256, 0, st>>>(kps, kun, n, count, depth, w, h, pitch, bf, uR, dep);
    return hipGetLastError();
}
