// orb_color.hip -- colour input for the front end: cvtColor to gray as
// Tracking::GrabImageStereo / GrabImageMonocular call it (src/Tracking.cc
// :250-276: CV_RGB2GRAY, CV_BGR2GRAY, CV_RGBA2GRAY, CV_BGRA2GRAY).
//
// OpenCV (2.4/3.x, absent here) converts 8-bit colour with 14-bit fixed-point
// weights: gray = (R*4899 + G*9617 + B*1868 + (1 << 13)) >> 14, the
// coefficients being round(16384 * {0.299, 0.587, 0.114}), which sum to 16384.
#include <hip/hip_runtime.h>
#include <stdint.h>

// One thread per 4 output pixels of a row; scn = 3 or 4 input channels,
// bidx = 0 when blue is channel 0 (BGR/BGRA), 2 when red is (RGB/RGBA).
__global__ __launch_bounds__(256) void gray_kernel(const uint8_t* __restrict__ src, size_t sstride, int scn, int bidx,
                                                   int w, int h, uint8_t* __restrict__ dst, size_t dstride)
{
    const int gx = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    const int x0 = gx * 4;
    if (x0 >= w || y >= h) return;
    const uint8_t* s = src + (size_t)y * sstride + (size_t)x0 * scn;
    unsigned out = 0;
    const int n = min(4, w - x0);
    for (int k = 0; k < n; k++) {
        const unsigned b = s[k * scn + bidx], g = s[k * scn + 1], r = s[k * scn + (bidx ^ 2)];
        const unsigned v = (b * 1868u + g * 9617u + r * 4899u + (1u << 13)) >> 14;
        out |= v << (8 * k);
    }
    uint8_t* d = dst + (size_t)y * dstride + x0;
    if (n == 4 && ((uintptr_t)d & 3) == 0) {
        *(unsigned*)d = out;
    } else {
        for (int k = 0; k < n; k++) d[k] = (uint8_t)(out >> (8 * k));
    }
}

extern "C" hipError_t orb_launch_gray(const uint8_t* src, size_t sstride, int scn, int bidx, int w, int h,
                                      uint8_t* dst, size_t dstride, hipStream_t st)
{
    if (w <= 0 || h <= 0) return hipSuccess;
    const dim3 grid((unsigned)((w + 1023) / 1024), (unsigned)h);
    gray_kernel<<<grid, 256, 0, st>>>(src, sstride, scn, bidx, w, h, dst, dstride);
    return hipGetLastError();
}

// remap(src, dst, mapx, mapy, INTER_LINEAR) for 8-bit images with the
// default constant zero border (Examples/Stereo/stereo_euroc.cc:173-180),
// as OpenCV 2.4/3.x computes it: X = cvRound(mapx*32), Y = cvRound(mapy*32),
// taps at (X >> 5, Y >> 5), weights of the 32 x 32 table (1-fx)(1-fy),
// fx(1-fy), (1-fx)fy, fx fy times 2^15 (exact integers: 32 * products of
// 5-bit fractions), out = (sum + 2^14) >> 15; taps outside the source read 0
// and a pixel whose whole 2x2 footprint is outside is 0.  One thread per 4
// output pixels of a row.
__device__ __forceinline__ int remap_px(const uint8_t* __restrict__ src, size_t sstride, int sw, int sh, float mx,
                                        float my)
{
    const float fx = fminf(fmaxf(mx * 32.f, -2147483648.f), 2147483520.f);
    const float fy = fminf(fmaxf(my * 32.f, -2147483648.f), 2147483520.f);
    const int X = __float2int_rn(fx), Y = __float2int_rn(fy);
    const int sx = min(max(X >> 5, -32768), 32767), sy = min(max(Y >> 5, -32768), 32767);
    if (sx >= sw || sx + 1 < 0 || sy >= sh || sy + 1 < 0) return 0;
    const int ax = X & 31, ay = Y & 31;
    const int w00 = (32 - ay) * (32 - ax) * 32, w01 = (32 - ay) * ax * 32;
    const int w10 = ay * (32 - ax) * 32, w11 = ay * ax * 32;
    const bool x0 = sx >= 0, x1 = sx + 1 < sw, y0 = sy >= 0, y1 = sy + 1 < sh;
    const uint8_t* r0 = src + (size_t)sy * sstride;
    const uint8_t* r1 = r0 + sstride;
    const int v00 = (x0 && y0) ? r0[sx] : 0, v01 = (x1 && y0) ? r0[sx + 1] : 0;
    const int v10 = (x0 && y1) ? r1[sx] : 0, v11 = (x1 && y1) ? r1[sx + 1] : 0;
    const int r = (v00 * w00 + v01 * w01 + v10 * w10 + v11 * w11 + (1 << 14)) >> 15;
    return min(max(r, 0), 255);
}

__global__ __launch_bounds__(256) void remap_kernel(const uint8_t* __restrict__ src, size_t sstride, int sw, int sh,
                                                    const float* __restrict__ mapx, const float* __restrict__ mapy,
                                                    int w, int h, uint8_t* __restrict__ dst, size_t dstride)
{
    const int x0 = (blockIdx.x * 256 + threadIdx.x) * 4;
    const int y = blockIdx.y;
    if (x0 >= w || y >= h) return;
    const float* mxr = mapx + (size_t)y * w;
    const float* myr = mapy + (size_t)y * w;
    const int n = min(4, w - x0);
    unsigned out = 0;
    for (int k = 0; k < n; k++) out |= (unsigned)remap_px(src, sstride, sw, sh, mxr[x0 + k], myr[x0 + k]) << (8 * k);
    uint8_t* d = dst + (size_t)y * dstride + x0;
    if (n == 4 && ((uintptr_t)d & 3) == 0) {
        *(unsigned*)d = out;
    } else {
        for (int k = 0; k < n; k++) d[k] = (uint8_t)(out >> (8 * k));
    }
}

extern "C" hipError_t orb_launch_remap(const uint8_t* src, size_t sstride, int sw, int sh, const float* mapx,
                                       const float* mapy, int w, int h, uint8_t* dst, size_t dstride, hipStream_t st)
{
    if (w <= 0 || h <= 0) return hipSuccess;
    const dim3 grid((unsigned)((w + 1023) / 1024), (unsigned)h);
    remap_kernel<<<grid, 256, 0, st>>>(src, sstride, sw, sh, mapx, mapy, w, h, dst, dstride);
    return hipGetLastError();
}
