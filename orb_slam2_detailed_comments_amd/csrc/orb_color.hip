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
