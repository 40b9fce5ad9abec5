// orb_rectify.cpp -- C ABI of stereo rectification in front of the
// pyramid (Examples/Stereo/stereo_euroc.cc:124-188): the maps of
// cv::initUndistortRectifyMap(K, D, R, P, size, CV_32FC1) computed once on the
// host in double with OpenCV 2.4/3.x's statements (or taken from the caller),
// kept in HBM, and cv::remap(INTER_LINEAR) per frame on the device, alone or
// straight into an extractor's input.
#include <cstring>
#include <vector>

#include "orb_internal.h"

using namespace orb_detail;

extern "C" hipError_t orb_launch_remap(const uint8_t* src, size_t sstride, int sw, int sh, const float* mapx,
                                       const float* mapy, int w, int h, uint8_t* dst, size_t dstride, hipStream_t st);

struct orb_rectify {
    CtxRef ctx;
    int w = 0, h = 0;
    float* d_mapx = nullptr;
    float* d_mapy = nullptr;
    uint8_t* d_src = nullptr;     // staged source frame
    size_t src_cap = 0;
    uint8_t* d_dst = nullptr;     // remapped frame (standalone calls)
    size_t dst_cap = 0;

    ~orb_rectify()
    {
        void* ps[] = {d_mapx, d_mapy, d_src, d_dst};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }

    int upload_maps(const float* mx, const float* my)
    {
        const size_t bytes = (size_t)w * h * sizeof(float);
        if (hipMalloc(&d_mapx, bytes) != hipSuccess || hipMalloc(&d_mapy, bytes) != hipSuccess) return ORB_E_OOM;
        HIP_TRY(hipMemcpy(d_mapx, mx, bytes, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_mapy, my, bytes, hipMemcpyHostToDevice));
        return ORB_OK;
    }

    int stage_src(const uint8_t* img, int sw, int sh, size_t stride)
    {
        int rc = ctx->grow((void**)&d_src, &src_cap, (size_t)sw * sh);
        if (rc) return rc;
        if (stride == (size_t)sw)
            HIP_TRY(hipMemcpyAsync(d_src, img, (size_t)sw * sh, hipMemcpyHostToDevice, ctx->stream));
        else
            HIP_TRY(hipMemcpy2DAsync(d_src, sw, img, stride, sw, sh, hipMemcpyHostToDevice, ctx->stream));
        return ORB_OK;
    }
};

namespace {

// Matx33d::inv(DECOMP_LU) of a 3x3 matrix: adjugate over the determinant.
bool invert3(const double a[9], double b[9])
{
    const double d = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[1] * (a[3] * a[8] - a[6] * a[5]) +
                     a[2] * (a[3] * a[7] - a[6] * a[4]);
    if (d == 0) return false;
    const double id = 1 / d;
    b[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    b[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    b[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    b[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    b[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    b[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    b[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    b[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    b[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return true;
}

}  // namespace

extern "C" {

// initUndistortRectifyMap(K, D, R, P, (width, height), CV_32FC1): for output
// pixel (j, i), the ray iR * (j, i, 1) with iR = (P R)^-1 accumulated along
// the row in double, the radial (k1..k6), tangential (p1, p2) and thin-prism
// (s1..s4) model, then (float)(fx * xd + u0), (float)(fy * yd + v0).
int orb_rectify_create(orb_context* ctx, const double* K, const double* D, int nD, const double* R, const double* P,
                       int width, int height, orb_rectify** out)
{
    if (!ctx || !K || !P || !out || width <= 0 || height <= 0) return ORB_E_ARG;
    if (!(nD == 0 || nD == 4 || nD == 5 || nD == 8 || nD == 12) || (nD && !D)) return ORB_E_ARG;
    *out = nullptr;
    double c[12] = {0};
    for (int i = 0; i < nD; i++) c[i] = D[i];
    const double k1 = c[0], k2 = c[1], p1 = c[2], p2 = c[3], k3 = c[4], k4 = c[5], k5 = c[6], k6 = c[7];
    const double s1 = c[8], s2 = c[9], s3 = c[10], s4 = c[11];
    double Rm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (R) std::memcpy(Rm, R, sizeof(Rm));
    double PR[9], ir[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) PR[i * 3 + j] = P[i * 3] * Rm[j] + P[i * 3 + 1] * Rm[3 + j] + P[i * 3 + 2] * Rm[6 + j];
    if (!invert3(PR, ir)) return ORB_E_ARG;
    const double u0 = K[2], v0 = K[5], fx = K[0], fy = K[4];
    std::vector<float> mx((size_t)width * height), my((size_t)width * height);
    for (int i = 0; i < height; i++) {
        double _x = i * ir[1] + ir[2], _y = i * ir[4] + ir[5], _w = i * ir[7] + ir[8];
        for (int j = 0; j < width; j++, _x += ir[0], _y += ir[3], _w += ir[6]) {
            const double w = 1. / _w, x = _x * w, y = _y * w;
            const double x2 = x * x, y2 = y * y;
            const double r2 = x2 + y2, _2xy = 2 * x * y;
            const double kr = (1 + ((k3 * r2 + k2) * r2 + k1) * r2) / (1 + ((k6 * r2 + k5) * r2 + k4) * r2);
            const double xd = (x * kr + p1 * _2xy + p2 * (r2 + 2 * x2) + s1 * r2 + s2 * r2 * r2);
            const double yd = (y * kr + p1 * (r2 + 2 * y2) + p2 * _2xy + s3 * r2 + s4 * r2 * r2);
            mx[(size_t)i * width + j] = (float)(fx * 1.0 * xd + u0);
            my[(size_t)i * width + j] = (float)(fy * 1.0 * yd + v0);
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    orb_rectify* r = new orb_rectify();
    r->ctx = ctx;
    r->w = width;
    r->h = height;
    const int rc = r->upload_maps(mx.data(), my.data());
    if (rc) {
        delete r;
        return rc;
    }
    *out = r;
    return ORB_OK;
}

int orb_rectify_from_maps(orb_context* ctx, const float* mapx, const float* mapy, int width, int height,
                          orb_rectify** out)
{
    if (!ctx || !mapx || !mapy || !out || width <= 0 || height <= 0) return ORB_E_ARG;
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    orb_rectify* r = new orb_rectify();
    r->ctx = ctx;
    r->w = width;
    r->h = height;
    const int rc = r->upload_maps(mapx, mapy);
    if (rc) {
        delete r;
        return rc;
    }
    *out = r;
    return ORB_OK;
}

void orb_rectify_destroy(orb_rectify* r)
{
    if (!r) return;
    (void)hipSetDevice(r->ctx->device);
    (void)hipStreamSynchronize(r->ctx->stream);
    delete r;
}

int orb_rectify_maps(orb_rectify* r, float* mapx, float* mapy)
{
    if (!r || !mapx || !mapy) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(hipStreamSynchronize(r->ctx->stream));
    HIP_TRY(hipMemcpy(mapx, r->d_mapx, (size_t)r->w * r->h * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(mapy, r->d_mapy, (size_t)r->w * r->h * 4, hipMemcpyDeviceToHost));
    return ORB_OK;
}

int orb_remap(orb_rectify* r, const uint8_t* src, int src_width, int src_height, size_t src_stride, uint8_t* dst,
              size_t dst_stride)
{
    if (!r || !src || !dst || src_width <= 0 || src_height <= 0 || src_stride < (size_t)src_width ||
        dst_stride < (size_t)r->w)
        return ORB_E_ARG;
    HIP_TRY(hipSetDevice(r->ctx->device));
    int rc = r->stage_src(src, src_width, src_height, src_stride);
    if (!rc) rc = r->ctx->grow((void**)&r->d_dst, &r->dst_cap, (size_t)r->w * r->h);
    if (rc) return rc;
    hipStream_t st = r->ctx->stream;
    HIP_TRY(orb_launch_remap(r->d_src, src_width, src_width, src_height, r->d_mapx, r->d_mapy, r->w, r->h, r->d_dst,
                             r->w, st));
    if (dst_stride == (size_t)r->w)   // one linear copy (a pitched one into pageable memory goes row by row)
        HIP_TRY(hipMemcpyAsync(dst, r->d_dst, (size_t)r->w * r->h, hipMemcpyDeviceToHost, st));
    else
        HIP_TRY(hipMemcpy2DAsync(dst, dst_stride, r->d_dst, r->w, r->w, r->h, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

int orb_extract_rectified(orb_extractor* ex, orb_rectify* r, const uint8_t* src, int src_width, int src_height,
                          size_t src_stride, orb_keypoint* keypoints, int capacity, uint8_t* descriptors, int* n)
{
    if (!ex || !r || !n || !src || src_width <= 0 || src_height <= 0 || src_stride < (size_t)src_width) return ORB_E_ARG;
    if (ex->ctx != r->ctx) return ORB_E_ARG;
    *n = 0;
    HIP_TRY(hipSetDevice(ex->ctx->device));
    int rc = ex->ensure(r->w, r->h);
    if (!rc) rc = r->stage_src(src, src_width, src_height, src_stride);
    if (rc) return rc;
    hipStream_t st = ex->ctx->stream;
    HIP_TRY(orb_launch_remap(r->d_src, src_width, src_width, src_height, r->d_mapx, r->d_mapy, r->w, r->h,
                             const_cast<uint8_t*>(ex->hjob.src), (size_t)ex->hjob.src_pitch, st));
    HIP_TRY(orb_launch_extract(ex->d_job, ex->d_rt, ex->h_rt, st, nullptr));
    return ex->read_results(keypoints, capacity, descriptors, n);
}

}  // extern "C"
