// orb_runtime.cpp -- host side of the C ABI declared in include/orb_gpu.h.
//
// The host work is the scalar bookkeeping that ORBextractor does on the CPU
// before any pixel is touched: scale tables and per-level quotas (the
// constructor), level sizes, the FAST cell grid and the quadtree roots of each
// level.  Those values are computed here with the same float statements as the
// reference and packed into ImageJob descriptors; the pixel and keypoint work
// runs in the gfx950 kernels of orb_kernels.hip.
#include "orb_internal.h"

// ====================================================================== C ABI
extern "C" {

int orb_context_create(int device, orb_context** out)
{
    if (!out) return ORB_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return ORB_E_DEVICE;
    HIP_TRY(hipSetDevice(device));
    orb_context* c = new orb_context();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ORB_E_DEVICE;
    }
    int umax[16], gauss[7];
    disc_umax(umax);
    gaussian_taps(gauss);
    if (orb_upload_constants(umax, gauss) != hipSuccess || orb_knn_scratch_setup(device) != 0) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return ORB_E_DEVICE;
    }
    *out = c;
    return ORB_OK;
}

void orb_context_destroy(orb_context* ctx) { orb_context_release_ref(ctx); }
}  // extern "C"

void orb_context_release_ref(orb_context* ctx)
{
    if (!ctx || ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_q) (void)hipFree(ctx->d_q);
    if (ctx->d_t) (void)hipFree(ctx->d_t);
    if (ctx->d_out) (void)hipFree(ctx->d_out);
    if (ctx->d_part) (void)hipFree(ctx->d_part);
    if (ctx->d_kq) (void)hipFree(ctx->d_kq);
    if (ctx->d_km) (void)hipFree(ctx->d_km);
    if (ctx->d_kout) (void)hipFree(ctx->d_kout);
    void* mp[] = {ctx->d_mp, ctx->d_mpo, ctx->d_mpb, ctx->d_mpd, ctx->d_color, ctx->d_win, ctx->d_wcand, ctx->d_fr};
    for (void* q : mp)
        if (q) (void)hipFree(q);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" {

int orb_context_synchronize(orb_context* ctx)
{
    if (!ctx) return ORB_E_ARG;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ORB_OK;
}

int orb_extractor_create(orb_context* ctx, int nfeatures, float scaleFactor, int nlevels, int iniThFAST,
                         int minThFAST, orb_extractor** out)
{
    if (!ctx || !out || nfeatures < 0 || nlevels < 1 || nlevels > ORB_MAX_LEVELS || !(scaleFactor > 1.0f))
        return ORB_E_ARG;
    orb_extractor* e = new orb_extractor();
    e->ctx = ctx;
    e->nfeatures = nfeatures;
    e->nlevels = nlevels;
    e->iniTh = iniThFAST;
    e->minTh = minThFAST;
    e->scaleFactorF = scaleFactor;
    e->scaleFactor = scaleFactor;
    e->init_tables();
    *out = e;
    return ORB_OK;
}

void orb_extractor_destroy(orb_extractor* ex)
{
    if (!ex) return;
    (void)hipSetDevice(ex->ctx->device);
    (void)hipStreamSynchronize(ex->ctx->stream);
    delete ex;
}

int orb_extractor_set_opencv_rounding(orb_extractor* ex, int mode)
{
    if (!ex || (mode != 0 && mode != 1)) return ORB_E_ARG;
    if (ex->rounding != mode) {
        ex->rounding = mode;
        ex->plan.W = ex->plan.H = 0;   // the next call rebuilds the level tables
    }
    return ORB_OK;
}

int orb_extractor_set_tile_heights(orb_extractor* ex, int n, const int* tile_h)
{
    if (!ex || n < 0 || n > ex->nlevels || (n && !tile_h)) return ORB_E_ARG;
    for (int l = 0; l < n; l++)
        if (tile_h[l] != 16 && tile_h[l] != 32) return ORB_E_ARG;
    bool change = false;
    for (int l = 0; l < n; l++) {
        change |= ex->tile_h[l] != tile_h[l];
        ex->tile_h[l] = tile_h[l];
    }
    if (change) ex->plan.W = ex->plan.H = 0;   // the next call re-plans the tiles
    return ORB_OK;
}

int orb_extractor_levels(const orb_extractor* ex) { return ex ? ex->nlevels : ORB_E_ARG; }

float orb_extractor_scale_factor(const orb_extractor* ex) { return ex ? (float)ex->scaleFactor : 0.f; }

int orb_extractor_scales(const orb_extractor* ex, float* scale, float* inv_scale, float* sigma2, float* inv_sigma2)
{
    if (!ex) return ORB_E_ARG;
    for (int i = 0; i < ex->nlevels; i++) {
        if (scale) scale[i] = ex->scale[i];
        if (inv_scale) inv_scale[i] = ex->invScale[i];
        if (sigma2) sigma2[i] = ex->sigma2[i];
        if (inv_sigma2) inv_sigma2[i] = ex->invSigma2[i];
    }
    return ORB_OK;
}

int orb_extractor_quotas(const orb_extractor* ex, int* quota)
{
    if (!ex || !quota) return ORB_E_ARG;
    for (int i = 0; i < ex->nlevels; i++) quota[i] = ex->quota[i];
    return ORB_OK;
}

int orb_extractor_capacity(const orb_extractor* ex, int width, int height)
{
    if (!ex) return ORB_E_ARG;
    Plan p;
    const int rc = ex->make_plan(width, height, p);
    return rc ? rc : p.out_cap;
}

int orb_extract_device(orb_extractor* ex, const uint8_t* d_image, int width, int height, size_t pitch)
{
    if (!ex || width < 0 || height < 0) return ORB_E_ARG;
    if (width == 0 || height == 0) return ORB_OK;
    if (!d_image || pitch < (size_t)width) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ex->ctx->device));
    int rc = ex->ensure(width, height);
    if (rc) return rc;
    rc = ex->upload_image(d_image, pitch, hipMemcpyDeviceToDevice);
    if (rc) return rc;
    HIP_TRY(orb_launch_extract(ex->d_job, ex->d_rt, ex->h_rt, ex->ctx->stream, nullptr));
    return ORB_OK;
}

int orb_extractor_results(orb_extractor* ex, orb_keypoint* kps, int cap, uint8_t* desc, int* n)
{
    if (!ex || !n) return ORB_E_ARG;
    *n = 0;
    if (!ex->d_mem) return ORB_OK;
    HIP_TRY(hipSetDevice(ex->ctx->device));
    return ex->read_results(kps, cap, desc, n);
}

int orb_extract(orb_extractor* ex, const uint8_t* image, int width, int height, size_t stride,
                orb_keypoint* keypoints, int capacity, uint8_t* descriptors, int* n)
{
    if (!ex || !n || width < 0 || height < 0) return ORB_E_ARG;
    *n = 0;
    if (width == 0 || height == 0) return ORB_OK;  // empty image: silent return (:1537)
    if (!image || stride < (size_t)width) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ex->ctx->device));
    int rc = ex->ensure(width, height);
    if (rc) return rc;
    rc = ex->upload_image(image, stride, hipMemcpyHostToDevice);
    if (rc) return rc;
    HIP_TRY(orb_launch_extract(ex->d_job, ex->d_rt, ex->h_rt, ex->ctx->stream, nullptr));
    return ex->read_results(keypoints, capacity, descriptors, n);
}

}  // extern "C"

// Colour rows -> the gray staging buffer of an extractor (or any pitched
// device destination), through the context's colour staging buffer.
static int upload_color(orb_context* ctx, const uint8_t* image, int width, int height, size_t stride, int channels,
                        int rgb, uint8_t* d_dst, size_t dst_pitch)
{
    const size_t row = (size_t)width * channels;
    int rc = ctx->grow((void**)&ctx->d_color, &ctx->color_cap, row * height);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    if (stride == row)
        HIP_TRY(hipMemcpyAsync(ctx->d_color, image, row * height, hipMemcpyHostToDevice, st));
    else
        HIP_TRY(hipMemcpy2DAsync(ctx->d_color, row, image, stride, row, height, hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_gray(ctx->d_color, row, channels, rgb ? 2 : 0, width, height, d_dst, dst_pitch, st));
    return ORB_OK;
}

extern "C" {

int orb_extract_color(orb_extractor* ex, const uint8_t* image, int width, int height, size_t stride, int channels,
                      int rgb, orb_keypoint* keypoints, int capacity, uint8_t* descriptors, int* n)
{
    if (!ex || !n || width < 0 || height < 0 || (channels != 3 && channels != 4)) return ORB_E_ARG;
    *n = 0;
    if (width == 0 || height == 0) return ORB_OK;
    if (!image || stride < (size_t)width * channels) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ex->ctx->device));
    int rc = ex->ensure(width, height);
    if (rc) return rc;
    rc = upload_color(ex->ctx, image, width, height, stride, channels, rgb,
                      const_cast<uint8_t*>(ex->hjob.src), (size_t)ex->hjob.src_pitch);
    if (rc) return rc;
    HIP_TRY(orb_launch_extract(ex->d_job, ex->d_rt, ex->h_rt, ex->ctx->stream, nullptr));
    return ex->read_results(keypoints, capacity, descriptors, n);
}

int orb_cvt_color_gray(orb_context* ctx, const uint8_t* image, int width, int height, size_t stride, int channels,
                       int rgb, uint8_t* gray, size_t gray_stride)
{
    if (!ctx || width < 0 || height < 0 || (channels != 3 && channels != 4)) return ORB_E_ARG;
    if (width == 0 || height == 0) return ORB_OK;
    if (!image || !gray || stride < (size_t)width * channels || gray_stride < (size_t)width) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ctx->grow(&ctx->d_fr, &ctx->fr_cap, (size_t)width * height);   // the context's scratch
    if (rc) return rc;
    uint8_t* d_gray = (uint8_t*)ctx->d_fr;
    rc = upload_color(ctx, image, width, height, stride, channels, rgb, d_gray, (size_t)width);
    if (rc) return rc;
    // rows at the host's own pitch go back in one linear copy (a pitched copy
    // into pageable memory goes row by row)
    if (gray_stride == (size_t)width)
        HIP_TRY(hipMemcpyAsync(gray, d_gray, (size_t)width * height, hipMemcpyDeviceToHost, ctx->stream));
    else
        HIP_TRY(hipMemcpy2DAsync(gray, gray_stride, d_gray, width, width, height, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ORB_OK;
}

int orb_extractor_pyramid_level(orb_extractor* ex, int level, uint8_t* dst, size_t dst_stride, int* width,
                                int* height)
{
    if (!ex || !ex->d_mem || level < 0 || level >= ex->nlevels) return ORB_E_ARG;
    if (width) *width = ex->plan.lv[level].w;
    if (height) *height = ex->plan.lv[level].h;
    return orb_extractor_debug_image(ex, level, 0, dst, dst_stride);
}

int orb_descriptor_distance(const uint8_t* a, const uint8_t* b)
{
    if (!a || !b) return ORB_E_ARG;
    int d = 0;
    for (int i = 0; i < 4; i++) {
        uint64_t x, y;
        std::memcpy(&x, a + 8 * i, 8);
        std::memcpy(&y, b + 8 * i, 8);
        d += __builtin_popcountll(x ^ y);
    }
    return d;
}

int orb_match_bf(orb_context* ctx, const uint8_t* query, int nq, const uint8_t* train, int nt, int th_low,
                 float ratio, int32_t* idx, int32_t* dist)
{
    if (!ctx || nq < 0 || nt < 0 || (nq && (!query || !idx || !dist)) || (nt && !train)) return ORB_E_ARG;
    if (nq == 0) return ORB_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    if ((size_t)nq > ctx->q_cap) {
        if (ctx->d_q) (void)hipFree(ctx->d_q);
        if (ctx->d_out) (void)hipFree(ctx->d_out);
        ctx->d_q = nullptr;
        ctx->d_out = nullptr;
        ctx->q_cap = 0;
        if (hipMalloc(&ctx->d_q, (size_t)nq * 32) != hipSuccess) return ORB_E_OOM;
        if (hipMalloc(&ctx->d_out, (size_t)nq * 8 + 64) != hipSuccess) return ORB_E_OOM;
        ctx->q_cap = nq;
    }
    if ((size_t)std::max(nt, 1) > ctx->t_cap) {
        if (ctx->d_t) (void)hipFree(ctx->d_t);
        ctx->d_t = nullptr;
        ctx->t_cap = 0;
        if (hipMalloc(&ctx->d_t, (size_t)std::max(nt, 1) * 32) != hipSuccess) return ORB_E_OOM;
        ctx->t_cap = std::max(nt, 1);
    }
    int* d_idx = ctx->d_out;
    int* d_dist = ctx->d_out + nq;
    int* d_n = ctx->d_out + 2 * (size_t)nq;
    const int nn[2] = {nq, nt};
    HIP_TRY(hipMemcpyAsync(ctx->d_q, query, (size_t)nq * 32, hipMemcpyHostToDevice, ctx->stream));
    if (nt) HIP_TRY(hipMemcpyAsync(ctx->d_t, train, (size_t)nt * 32, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_n, nn, sizeof(nn), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(orb_launch_match_bf(ctx->d_q, d_n, ctx->d_t, d_n + 1, th_low, ratio, d_idx, d_dist, nq, ctx->stream));
    HIP_TRY(hipMemcpyAsync(idx, d_idx, (size_t)nq * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dist, d_dist, (size_t)nq * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ORB_OK;
}

int orb_stereo_match(orb_extractor* left, orb_extractor* right, const orb_keypoint* kl, const uint8_t* dl, int nl,
                     const orb_keypoint* kr, const uint8_t* dr, int nr, float bf, float b, float* uRight,
                     float* depth)
{
    if (!left || !right || nl < 0 || nr < 0 || (nl && (!kl || !dl || !uRight || !depth)) || (nr && (!kr || !dr)))
        return ORB_E_ARG;
    if (nl == 0) return ORB_OK;
    // both pyramids are read on the left context's stream, so the right one
    // must have been written on that stream too
    if (!left->d_mem || !right->d_mem || left->plan.W != right->plan.W || left->plan.H != right->plan.H ||
        left->nlevels != right->nlevels || left->ctx != right->ctx)
        return ORB_E_ARG;
    for (int i = 0; i < nl; i++)
        if (kl[i].octave < 0 || kl[i].octave >= left->nlevels) return ORB_E_ARG;
    for (int i = 0; i < nr; i++)
        if (kr[i].octave < 0 || kr[i].octave >= left->nlevels) return ORB_E_ARG;
    // stereo scratch: right keypoints sorted in LDS (<= 4096) and the
    // left extractor's per-keypoint arrays
    if (nl > left->plan.out_cap || nr > std::min(left->plan.out_cap, right->plan.out_cap) || nl > 4096 ||
        nr > 4096)
        return ORB_E_CAPACITY;
    HIP_TRY(hipSetDevice(left->ctx->device));
    hipStream_t st = left->ctx->stream;
    int rc = left->ensure_pair(right);
    if (rc) return rc;
    rc = left->ensure_inputs(std::max(std::max(nl, nr), 1));
    if (rc) return rc;
    OrbKeyPoint* d_kl = left->d_kin;
    OrbKeyPoint* d_kr = left->d_kin + left->in_cap;
    uint8_t* d_dl = left->d_din;
    uint8_t* d_dr = left->d_din + (size_t)left->in_cap * 32;
    static_assert(sizeof(orb_keypoint) == sizeof(OrbKeyPoint), "keypoint layout");
    HIP_TRY(hipMemcpyAsync(d_kl, kl, (size_t)nl * sizeof(OrbKeyPoint), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_dl, dl, (size_t)nl * 32, hipMemcpyHostToDevice, st));
    if (nr) {
        HIP_TRY(hipMemcpyAsync(d_kr, kr, (size_t)nr * sizeof(OrbKeyPoint), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_dr, dr, (size_t)nr * 32, hipMemcpyHostToDevice, st));
    }
    const int counts[2] = {nl, nr};
    HIP_TRY(hipMemcpyAsync(left->d_misc + 1, counts, sizeof(counts), hipMemcpyHostToDevice, st));
    StereoArgs a;
    left->stereo_args(a, right, left->d_pair_jobs, left->d_pair_jobs + 1, bf, b);
    a.kl = d_kl;
    a.dl = d_dl;
    a.nl = left->d_misc + 1;
    a.kr = d_kr;
    a.dr = d_dr;
    a.nr = left->d_misc + 2;
    a.cap_l = nl;
    rc = left->upload_stereo_args(1, a);
    if (rc) return rc;
    HIP_TRY(orb_launch_stereo(left->d_sargs + 1, 1, nl, st, nullptr));
    HIP_TRY(hipMemcpyAsync(uRight, left->d_uR, (size_t)nl * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(depth, left->d_depth, (size_t)nl * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

int orb_stereo_frame(orb_extractor* left, orb_extractor* right, const uint8_t* im_left, const uint8_t* im_right,
                     int width, int height, size_t stride, float bf, float b, orb_keypoint* kl, uint8_t* dl,
                     int* nl, orb_keypoint* kr, uint8_t* dr, int* nr, int capacity, float* uRight, float* depth)
{
    if (!left || !right || left == right || !nl || !nr || width < 0 || height < 0) return ORB_E_ARG;
    *nl = *nr = 0;
    if (width == 0 || height == 0) return ORB_OK;
    if (!im_left || !im_right || stride < (size_t)width || left->nlevels != right->nlevels ||
        left->ctx != right->ctx)
        return ORB_E_ARG;
    HIP_TRY(hipSetDevice(left->ctx->device));
    hipStream_t st = left->ctx->stream;
    int rc = left->ensure(width, height);
    if (!rc) rc = right->ensure(width, height);
    if (!rc) rc = left->ensure_pair(right);
    if (rc) return rc;
    if (left->plan.out_cap > 4096 || right->plan.out_cap > 4096) return ORB_E_ARG;
    rc = left->upload_image(im_left, stride, hipMemcpyHostToDevice);
    if (!rc) rc = right->upload_image(im_right, stride, hipMemcpyHostToDevice);
    if (rc) return rc;
    HIP_TRY(orb_launch_extract(left->d_pair_jobs, left->d_pair_rt, left->h_pair_rt, st, nullptr));
    StereoArgs a;
    left->stereo_args(a, right, left->d_pair_jobs, left->d_pair_jobs + 1, bf, b);
    a.kl = left->hjob.kps;
    a.dl = left->hjob.desc;
    a.nl = left->hjob.counters + ORB_CNT_NOUT;
    a.kr = right->hjob.kps;
    a.dr = right->hjob.desc;
    a.nr = right->hjob.counters + ORB_CNT_NOUT;
    a.cap_l = left->plan.out_cap;
    rc = left->upload_stereo_args(0, a);
    if (rc) return rc;
    HIP_TRY(orb_launch_stereo(left->d_sargs, 1, a.cap_l, st, nullptr));
    rc = left->read_results(kl, capacity, dl, nl);
    if (rc) return rc;
    rc = right->read_results(kr, capacity, dr, nr);
    if (rc) return rc;
    if (*nl > 0) {
        if (uRight) HIP_TRY(hipMemcpyAsync(uRight, left->d_uR, (size_t)*nl * 4, hipMemcpyDeviceToHost, st));
        if (depth) HIP_TRY(hipMemcpyAsync(depth, left->d_depth, (size_t)*nl * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return ORB_OK;
}

int orb_extractor_debug_image(orb_extractor* ex, int level, int kind, uint8_t* dst, size_t dst_stride)
{
    if (!ex || !ex->d_mem || level < 0 || level >= ex->nlevels || kind < 0 || kind > 1) return ORB_E_ARG;
    if (!dst) return ORB_OK;
    const LevelJob& L = ex->hjob.lv[level];
    const uint8_t* src = kind == 0 ? L.lvl : L.blur;
    if (dst_stride < (size_t)L.w) return ORB_E_ARG;
    orb_context* ctx = ex->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (dst_stride == (size_t)L.w) {
        // rows packed on the device first, then one linear copy (a pitched
        // copy into pageable memory goes row by row)
        const int rc = ctx->grow(&ctx->d_fr, &ctx->fr_cap, (size_t)L.w * L.h);
        if (rc) return rc;
        HIP_TRY(hipMemcpy2DAsync(ctx->d_fr, L.w, src, L.pitch, L.w, L.h, hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(dst, ctx->d_fr, (size_t)L.w * L.h, hipMemcpyDeviceToHost, st));
    } else {
        HIP_TRY(hipMemcpy2DAsync(dst, dst_stride, src, L.pitch, L.w, L.h, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

int orb_extractor_debug_candidates(orb_extractor* ex, int level, int* xyr, int capacity, int* n)
{
    if (!ex || !ex->d_mem || !n || level < 0 || level >= ex->nlevels) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ex->ctx->device));
    HIP_TRY(hipStreamSynchronize(ex->ctx->stream));
    int c[ORB_CNT_SLOTS];
    HIP_TRY(hipMemcpy(c, ex->hjob.counters, sizeof(c), hipMemcpyDeviceToHost));
    const LevelJob& L = ex->hjob.lv[level];
    int k = std::min(c[ORB_CNT_KEYS + level], L.key_cap);
    std::vector<unsigned long long> all(k);
    std::vector<uint8_t> flags(L.ncells);
    if (k) HIP_TRY(hipMemcpy(all.data(), L.keys, (size_t)k * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(flags.data(), L.cellflag, L.ncells, hipMemcpyDeviceToHost));
    std::vector<unsigned long long> keys;
    for (unsigned long long key : all) {
        const unsigned ok = (unsigned)key;
        const bool flagged = flags[((ok >> 20) & 255) * L.nCols + ((ok >> 12) & 255)] != 0;
        if ((key & ORB_KEY_RETRY) ? !flagged : flagged) keys.push_back(key & ~ORB_KEY_RETRY);
    }
    k = (int)keys.size();
    *n = k;
    if (k > capacity) return ORB_E_CAPACITY;
    std::sort(keys.begin(), keys.end(), [](unsigned long long a, unsigned long long b) {
        return (unsigned)a < (unsigned)b;
    });
    for (int i = 0; i < k; i++) {
        const unsigned ok = (unsigned)keys[i];
        const int xl = ok & 63, yl = (ok >> 6) & 63, j = (ok >> 12) & 255, r = (ok >> 20) & 255;
        xyr[3 * i] = j * L.wCell + xl;
        xyr[3 * i + 1] = r * L.hCell + yl;
        xyr[3 * i + 2] = (int)(keys[i] >> 32);
    }
    return ORB_OK;
}

}  // extern "C"

// ================================================================== batch
// A fixed set of image slots extracted together: one launch per stage over
// every slot (mode 0), the same launches replayed from one captured hipGraph
// (mode 1), or the per-image launch sequence for comparison (mode 2).
// Optional stereo pairs among the slots are matched in one multi-pair launch.
struct orb_batch {
    CtxRef ctx;
    std::vector<orb_extractor*> ex;
    ImageJob* d_jobs = nullptr;
    RunTable* h_rt = nullptr;
    RunTable* d_rt = nullptr;
    std::vector<int> pl, pr;
    std::vector<StereoArgs> h_sargs;
    StereoArgs* d_sargs = nullptr;
    int nwaves = 0;                // largest left capacity among the pairs
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    // Graph branches (mode 1): contiguous slot groups holding whole pairs.
    struct Branch {
        int s = 0, n = 0;                  // slots [s, s + n)
        std::unique_ptr<RunTable> h_rt;
        RunTable* d_rt = nullptr;
        StereoArgs* d_sargs = nullptr;
        int npairs = 0, nwaves = 0;
    };
    int want_branches = 1;
    std::vector<Branch> br;
    std::vector<hipStream_t> br_st;        // side streams, default priority
    std::vector<hipEvent_t> br_ev;         // fork + one join per side stream

    // ev (optional, mode 0): stage events, see orb_batch_stage_times.
    int enqueue(int mode, hipEvent_t* ev = nullptr)
    {
        hipStream_t st = ctx->stream;
        if (mode == 2) {
            for (orb_extractor* e : ex) HIP_TRY(orb_launch_extract(e->d_job, e->d_rt, e->h_rt, st, nullptr));
        } else {
            HIP_TRY(orb_launch_extract(d_jobs, d_rt, h_rt, st, ev));
        }
        if (!pl.empty())
            HIP_TRY(orb_launch_stereo(d_sargs, (int)pl.size(), nwaves, st, ev ? ev + h_rt->max_levels + 3 : nullptr));
        return ORB_OK;
    }

    // Cut the slots into up to want_branches groups of whole pairs.
    std::vector<int> cuts() const
    {
        const int n = (int)ex.size();
        std::vector<int> c{0};
        for (int k = 1; k < want_branches; k++) {
            const int target = (int)((long long)n * k / want_branches);
            int best = -1;
            for (int d = 0; d < n && best < 0; d++)
                for (int m : {target + d, target - d}) {
                    if (m <= c.back() || m >= n) continue;
                    bool ok = true;
                    for (size_t p = 0; p < pl.size() && ok; p++)
                        ok = (pl[p] < m) == (pr[p] < m);
                    if (ok) {
                        best = m;
                        break;
                    }
                }
            if (best > c.back()) c.push_back(best);
        }
        c.push_back(n);
        return c;
    }

    int build_branches()
    {
        drop_branches();
        const std::vector<int> c = cuts();
        if (c.size() <= 2) return ORB_OK;
        for (size_t g = 0; g + 1 < c.size(); g++) {
            Branch B;
            B.s = c[g];
            B.n = c[g + 1] - c[g];
            B.h_rt.reset(new RunTable());
            const int rc_t = orb_extractor::build_table(ex.data() + B.s, B.n, *B.h_rt);
            if (rc_t) return rc_t;
            std::vector<StereoArgs> sub;
            for (size_t p = 0; p < pl.size(); p++)
                if (pl[p] >= B.s && pl[p] < B.s + B.n) {
                    sub.push_back(h_sargs[p]);
                    B.nwaves = std::max(B.nwaves, h_sargs[p].cap_l);
                }
            B.npairs = (int)sub.size();
            br.push_back(std::move(B));
            Branch& R = br.back();
            if (hipMalloc(&R.d_rt, sizeof(RunTable)) != hipSuccess) return ORB_E_OOM;
            HIP_TRY(hipMemcpy(R.d_rt, R.h_rt.get(), sizeof(RunTable), hipMemcpyHostToDevice));
            if (R.npairs) {
                if (hipMalloc(&R.d_sargs, R.npairs * sizeof(StereoArgs)) != hipSuccess) return ORB_E_OOM;
                HIP_TRY(hipMemcpy(R.d_sargs, sub.data(), R.npairs * sizeof(StereoArgs), hipMemcpyHostToDevice));
            }
        }
        while (br_st.size() + 1 < br.size()) {
            hipStream_t s2 = nullptr;
            HIP_TRY(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            br_st.push_back(s2);
        }
        while (br_ev.size() < br.size()) {
            hipEvent_t e = nullptr;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            br_ev.push_back(e);
        }
        return ORB_OK;
    }

    // Captured body of mode 1 with branches: fork from the batch stream, one
    // group per stream, join back.
    int enqueue_branches()
    {
        hipStream_t st = ctx->stream;
        HIP_TRY(hipEventRecord(br_ev[0], st));
        for (size_t g = 1; g < br.size(); g++) HIP_TRY(hipStreamWaitEvent(br_st[g - 1], br_ev[0], 0));
        for (size_t g = 0; g < br.size(); g++) {
            const Branch& B = br[g];
            hipStream_t s = g == 0 ? st : br_st[g - 1];
            HIP_TRY(orb_launch_extract(d_jobs + B.s, B.d_rt, B.h_rt.get(), s, nullptr));
            if (B.npairs) HIP_TRY(orb_launch_stereo(B.d_sargs, B.npairs, B.nwaves, s, nullptr));
        }
        for (size_t g = 1; g < br.size(); g++) {
            HIP_TRY(hipEventRecord(br_ev[g], br_st[g - 1]));
            HIP_TRY(hipStreamWaitEvent(st, br_ev[g], 0));
        }
        return ORB_OK;
    }

    void drop_branches()
    {
        for (Branch& B : br) {
            if (B.d_rt) (void)hipFree(B.d_rt);
            if (B.d_sargs) (void)hipFree(B.d_sargs);
        }
        br.clear();
    }

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        drop_branches();
    }

    // pinned staging of orb_batch_read_all
    int* h_ctr = nullptr;
    char* h_res = nullptr;
    size_t h_res_cap = 0;

    // Pinned transfers (orb_batch_pinned_images / _upload_pinned /
    // _read_async).  Inputs: every slot's image at its device pitch in one
    // pinned buffer, one host-to-device copy into a device staging buffer,
    // and one copy-list launch into the slots' images.  Outputs: every slot's
    // counters, keypoints and descriptors and every pair's uRight / depth at
    // fixed offsets of one device buffer (one copy-list launch, valid entries
    // only), then one device-to-host copy into its pinned twin.
    uint8_t* h_in = nullptr;
    uint8_t* d_in = nullptr;
    std::vector<size_t> in_off;
    CopyOp* d_in_ops = nullptr;
    size_t in_bytes = 0;
    char* h_out = nullptr;
    char* d_out = nullptr;
    CopyOp* d_out_ops = nullptr;
    size_t out_bytes = 0;
    int n_out_ops = 0, out_pairs = -1;      // pair count the output layout was built for
    std::vector<size_t> out_ctr, out_kps, out_desc, out_ur, out_dep;
    hipEvent_t out_ev = nullptr;            // recorded after the device-to-host copy
    bool out_ready = false;                 // a read has been enqueued since the layout was built

    static size_t up256(size_t v) { return (v + 255) / 256 * 256; }

    int setup_pinned_in()
    {
        if (h_in) return ORB_OK;
        const int n = (int)ex.size();
        in_off.assign(n, 0);
        size_t o = 0;
        for (int s = 0; s < n; s++) {
            in_off[s] = o;
            o += up256((size_t)ex[s]->hjob.src_pitch * ex[s]->plan.H);
        }
        in_bytes = o;
        if (hipHostMalloc((void**)&h_in, in_bytes) != hipSuccess) {
            h_in = nullptr;
            return ORB_E_OOM;
        }
        std::memset(h_in, 0, in_bytes);
        if (hipMalloc((void**)&d_in, in_bytes) != hipSuccess || hipMalloc(&d_in_ops, n * sizeof(CopyOp)) != hipSuccess)
            return ORB_E_OOM;
        return write_in_ops();
    }

    // The staging-to-image copy list, from the slots' current image addresses.
    int write_in_ops()
    {
        const int n = (int)ex.size();
        std::vector<CopyOp> ops(n);
        for (int s = 0; s < n; s++) {
            const ImageJob& J = ex[s]->hjob;
            ops[s] = CopyOp{(const unsigned*)(d_in + in_off[s]), (unsigned*)const_cast<uint8_t*>(J.src), nullptr,
                            (int)((size_t)J.src_pitch * ex[s]->plan.H / 4), 0};
        }
        HIP_TRY(hipMemcpy(d_in_ops, ops.data(), n * sizeof(CopyOp), hipMemcpyHostToDevice));
        return ORB_OK;
    }

    // After the slots' plans or the stereo pairs change (stream idle): the
    // input copy list is rewritten and the output layout rebuilt on the next
    // read.
    int refresh_pinned()
    {
        drop_pinned_out();
        return d_in_ops ? write_in_ops() : ORB_OK;
    }

    void drop_pinned_out()
    {
        if (h_out) (void)hipHostFree(h_out);
        if (d_out) (void)hipFree(d_out);
        if (d_out_ops) (void)hipFree(d_out_ops);
        h_out = nullptr;
        d_out = nullptr;
        d_out_ops = nullptr;
        out_bytes = 0;
        n_out_ops = 0;
        out_pairs = -1;
        out_ready = false;
    }

    int setup_pinned_out()
    {
        if (h_out && out_pairs == (int)pl.size()) return ORB_OK;
        drop_pinned_out();
        const int n = (int)ex.size(), np = (int)pl.size();
        out_ctr.assign(n, 0);
        out_kps.assign(n, 0);
        out_desc.assign(n, 0);
        out_ur.assign(np, 0);
        out_dep.assign(np, 0);
        std::vector<CopyOp> ops;
        size_t o = 0;
        for (int s = 0; s < n; s++) {
            const ImageJob& J = ex[s]->hjob;
            const size_t cap = (size_t)J.out_cap;
            out_ctr[s] = o;
            o += up256(ORB_CNT_SLOTS * 4);
            out_kps[s] = o;
            o += up256(cap * sizeof(OrbKeyPoint));
            out_desc[s] = o;
            o += up256(cap * 32);
        }
        for (int p = 0; p < np; p++) {
            const size_t cap = (size_t)ex[pl[p]]->hjob.out_cap;
            out_ur[p] = o;
            o += up256(cap * 4);
            out_dep[p] = o;
            o += up256(cap * 4);
        }
        out_bytes = o;
        if (hipHostMalloc((void**)&h_out, out_bytes) != hipSuccess) {
            h_out = nullptr;
            return ORB_E_OOM;
        }
        if (hipMalloc((void**)&d_out, out_bytes) != hipSuccess) return ORB_E_OOM;
        for (int s = 0; s < n; s++) {
            const ImageJob& J = ex[s]->hjob;
            const int* cnt = J.counters + ORB_CNT_NOUT;
            ops.push_back(CopyOp{(const unsigned*)J.counters, (unsigned*)(d_out + out_ctr[s]), nullptr, ORB_CNT_SLOTS, 0});
            ops.push_back(CopyOp{(const unsigned*)J.kps, (unsigned*)(d_out + out_kps[s]), cnt,
                                 J.out_cap * (int)(sizeof(OrbKeyPoint) / 4), (int)(sizeof(OrbKeyPoint) / 4)});
            ops.push_back(CopyOp{(const unsigned*)J.desc, (unsigned*)(d_out + out_desc[s]), cnt, J.out_cap * 8, 8});
        }
        for (int p = 0; p < np; p++) {
            const orb_extractor* L = ex[pl[p]];
            const int* cnt = L->hjob.counters + ORB_CNT_NOUT;
            ops.push_back(CopyOp{(const unsigned*)L->d_uR, (unsigned*)(d_out + out_ur[p]), cnt, L->hjob.out_cap, 1});
            ops.push_back(CopyOp{(const unsigned*)L->d_depth, (unsigned*)(d_out + out_dep[p]), cnt, L->hjob.out_cap, 1});
        }
        n_out_ops = (int)ops.size();
        if (hipMalloc(&d_out_ops, ops.size() * sizeof(CopyOp)) != hipSuccess) return ORB_E_OOM;
        HIP_TRY(hipMemcpy(d_out_ops, ops.data(), ops.size() * sizeof(CopyOp), hipMemcpyHostToDevice));
        if (!out_ev) HIP_TRY(hipEventCreateWithFlags(&out_ev, hipEventDisableTiming));
        out_pairs = np;
        return ORB_OK;
    }

    // After the last read: wait for it, then check the slot's error bits.
    int pinned_count(int slot, int* n)
    {
        if (!h_out || !out_ready) return ORB_E_ARG;
        HIP_TRY(hipEventSynchronize(out_ev));
        const int* c = (const int*)(h_out + out_ctr[slot]);
        if (c[ORB_CNT_ERR]) return ORB_E_OVERFLOW;
        *n = c[ORB_CNT_NOUT];
        return ORB_OK;
    }

    ~orb_batch()
    {
        if (h_ctr) (void)hipHostFree(h_ctr);
        if (h_res) (void)hipHostFree(h_res);
        if (h_in) (void)hipHostFree(h_in);
        if (d_in) (void)hipFree(d_in);
        if (d_in_ops) (void)hipFree(d_in_ops);
        drop_pinned_out();
        if (out_ev) (void)hipEventDestroy(out_ev);
        drop_graph();
        for (hipStream_t s2 : br_st) (void)hipStreamDestroy(s2);
        for (hipEvent_t e : br_ev) (void)hipEventDestroy(e);
        for (orb_extractor* e : ex) delete e;
        if (d_jobs) (void)hipFree(d_jobs);
        if (d_rt) (void)hipFree(d_rt);
        if (h_rt) (void)hipHostFree(h_rt);
        if (d_sargs) (void)hipFree(d_sargs);
    }
};

extern "C" {

int orb_batch_create(orb_context* ctx, int n, const int* widths, const int* heights, const int* nfeatures,
                     float scaleFactor, int nlevels, int iniThFAST, int minThFAST, orb_batch** out)
{
    if (!ctx || !out || n <= 0 || n > ORB_MAX_JOBS || !widths || !heights || !nfeatures) return ORB_E_ARG;
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    orb_batch* b = new orb_batch();
    b->ctx = ctx;
    for (int i = 0; i < n; i++) {
        orb_extractor* e = nullptr;
        int rc = orb_extractor_create(ctx, nfeatures[i], scaleFactor, nlevels, iniThFAST, minThFAST, &e);
        if (!rc) {
            b->ex.push_back(e);
            rc = e->ensure(widths[i], heights[i]);
        }
        if (rc) {
            delete b;
            return rc;
        }
    }
    std::vector<ImageJob> jobs(n);
    for (int i = 0; i < n; i++) jobs[i] = b->ex[i]->hjob;
    if (hipMalloc(&b->d_jobs, n * sizeof(ImageJob)) != hipSuccess ||
        hipMalloc(&b->d_rt, sizeof(RunTable)) != hipSuccess ||
        hipHostMalloc(&b->h_rt, sizeof(RunTable)) != hipSuccess) {
        delete b;
        return ORB_E_OOM;
    }
    if (orb_extractor::build_table(b->ex.data(), n, *b->h_rt) != ORB_OK) {   // slots share tile heights
        delete b;
        return ORB_E_ARG;
    }
    if (hipMemcpy(b->d_jobs, jobs.data(), n * sizeof(ImageJob), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_rt, b->h_rt, sizeof(RunTable), hipMemcpyHostToDevice) != hipSuccess) {
        delete b;
        return ORB_E_DEVICE;
    }
    *out = b;
    return ORB_OK;
}

void orb_batch_destroy(orb_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

int orb_batch_size(const orb_batch* b) { return b ? (int)b->ex.size() : ORB_E_ARG; }

int orb_batch_image(orb_batch* b, int slot, uint8_t** d_image, size_t* pitch, int* width, int* height)
{
    if (!b || slot < 0 || slot >= (int)b->ex.size()) return ORB_E_ARG;
    const orb_extractor* e = b->ex[slot];
    if (d_image) *d_image = const_cast<uint8_t*>(e->hjob.src);
    if (pitch) *pitch = (size_t)e->hjob.src_pitch;
    if (width) *width = e->plan.W;
    if (height) *height = e->plan.H;
    return ORB_OK;
}

int orb_batch_upload(orb_batch* b, int slot, const uint8_t* image, size_t stride)
{
    if (!b || slot < 0 || slot >= (int)b->ex.size() || !image) return ORB_E_ARG;
    orb_extractor* e = b->ex[slot];
    if (stride < (size_t)e->plan.W) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    int rc = e->upload_image(image, stride, hipMemcpyHostToDevice);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return ORB_OK;
}

int orb_batch_upload_images(orb_batch* b, int n, const int* slots, const uint8_t* const* images,
                            const size_t* strides)
{
    if (!b || n < 0 || (n && (!slots || !images || !strides))) return ORB_E_ARG;
    for (int i = 0; i < n; i++)
        if (slots[i] < 0 || slots[i] >= (int)b->ex.size() || !images[i] || strides[i] < (size_t)b->ex[slots[i]]->plan.W)
            return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    for (int i = 0; i < n; i++) {
        const int rc = b->ex[slots[i]]->upload_image(images[i], strides[i], hipMemcpyHostToDevice);
        if (rc) return rc;
    }
    return ORB_OK;
}

int orb_batch_read_all(orb_batch* b, orb_keypoint* keypoints, uint8_t* descriptors, int capacity, int* counts,
                       float* uRight, float* depth)
{
    if (!b || capacity < 0 || !counts || (capacity && (!keypoints || !descriptors))) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const int n = (int)b->ex.size(), np = (int)b->pl.size();
    if (!b->h_ctr && hipHostMalloc((void**)&b->h_ctr, (size_t)n * ORB_CNT_SLOTS * sizeof(int)) != hipSuccess) {
        b->h_ctr = nullptr;
        return ORB_E_OOM;
    }
    for (int s = 0; s < n; s++)
        HIP_TRY(hipMemcpyAsync(b->h_ctr + (size_t)s * ORB_CNT_SLOTS, b->ex[s]->hjob.counters,
                               ORB_CNT_SLOTS * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    size_t need = 0;
    for (int s = 0; s < n; s++) {
        const int* c = b->h_ctr + (size_t)s * ORB_CNT_SLOTS;
        if (c[ORB_CNT_ERR]) return ORB_E_OVERFLOW;
        counts[s] = c[ORB_CNT_NOUT];
        if (counts[s] > capacity) return ORB_E_CAPACITY;
        need += (size_t)counts[s] * (sizeof(orb_keypoint) + 32);
    }
    for (int p = 0; p < np; p++) need += (size_t)counts[b->pl[p]] * 8;
    if (need > b->h_res_cap) {
        if (b->h_res) (void)hipHostFree(b->h_res);
        b->h_res = nullptr;
        b->h_res_cap = 0;
        if (hipHostMalloc((void**)&b->h_res, need) != hipSuccess) return ORB_E_OOM;
        b->h_res_cap = need;
    }
    size_t o = 0;
    std::vector<size_t> at(n + np);
    for (int s = 0; s < n; s++) {
        const size_t c = (size_t)counts[s];
        at[s] = o;
        if (c) {
            HIP_TRY(hipMemcpyAsync(b->h_res + o, b->ex[s]->hjob.kps, c * sizeof(orb_keypoint), hipMemcpyDeviceToHost,
                                   st));
            HIP_TRY(hipMemcpyAsync(b->h_res + o + c * sizeof(orb_keypoint), b->ex[s]->hjob.desc, c * 32,
                                   hipMemcpyDeviceToHost, st));
        }
        o += c * (sizeof(orb_keypoint) + 32);
    }
    for (int p = 0; p < np; p++) {
        const orb_extractor* L = b->ex[b->pl[p]];
        const size_t c = (size_t)counts[b->pl[p]];
        at[n + p] = o;
        if (c) {
            HIP_TRY(hipMemcpyAsync(b->h_res + o, L->d_uR, c * 4, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(b->h_res + o + c * 4, L->d_depth, c * 4, hipMemcpyDeviceToHost, st));
        }
        o += c * 8;
    }
    HIP_TRY(hipStreamSynchronize(st));
    for (int s = 0; s < n; s++) {
        const size_t c = (size_t)counts[s];
        std::memcpy(keypoints + (size_t)s * capacity, b->h_res + at[s], c * sizeof(orb_keypoint));
        std::memcpy(descriptors + (size_t)s * capacity * 32, b->h_res + at[s] + c * sizeof(orb_keypoint), c * 32);
    }
    for (int p = 0; p < np; p++) {
        const size_t c = (size_t)counts[b->pl[p]];
        if (uRight) std::memcpy(uRight + (size_t)p * capacity, b->h_res + at[n + p], c * 4);
        if (depth) std::memcpy(depth + (size_t)p * capacity, b->h_res + at[n + p] + c * 4, c * 4);
    }
    return ORB_OK;
}

int orb_batch_pinned_images(orb_batch* b, uint8_t** images, size_t* pitches)
{
    if (!b || !images) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    const int rc = b->setup_pinned_in();
    if (rc) return rc;
    for (size_t s = 0; s < b->ex.size(); s++) {
        images[s] = b->h_in + b->in_off[s];
        if (pitches) pitches[s] = (size_t)b->ex[s]->hjob.src_pitch;
    }
    return ORB_OK;
}

int orb_batch_upload_pinned(orb_batch* b)
{
    if (!b || !b->h_in) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    HIP_TRY(hipMemcpyAsync(b->d_in, b->h_in, b->in_bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_copy_ops(b->d_in_ops, (int)b->ex.size(), 64, st));
    return ORB_OK;
}

int orb_batch_read_async(orb_batch* b)
{
    if (!b) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    const int rc = b->setup_pinned_out();
    if (rc) return rc;
    hipStream_t st = b->ctx->stream;
    HIP_TRY(orb_launch_copy_ops(b->d_out_ops, b->n_out_ops, 8, st));
    HIP_TRY(hipMemcpyAsync(b->h_out, b->d_out, b->out_bytes, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(b->out_ev, st));
    b->out_ready = true;
    return ORB_OK;
}

int orb_batch_pinned_results(orb_batch* b, int slot, const orb_keypoint** keypoints, const uint8_t** descriptors,
                             int* n)
{
    if (!b || slot < 0 || slot >= (int)b->ex.size() || !n) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    const int rc = b->pinned_count(slot, n);
    if (rc) return rc;
    if (keypoints) *keypoints = (const orb_keypoint*)(b->h_out + b->out_kps[slot]);
    if (descriptors) *descriptors = (const uint8_t*)(b->h_out + b->out_desc[slot]);
    return ORB_OK;
}

int orb_batch_pinned_stereo(orb_batch* b, int pair, const float** uRight, const float** depth, int* n)
{
    if (!b || pair < 0 || pair >= (int)b->pl.size() || !n || b->out_pairs != (int)b->pl.size()) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    const int rc = b->pinned_count(b->pl[pair], n);
    if (rc) return rc;
    if (uRight) *uRight = (const float*)(b->h_out + b->out_ur[pair]);
    if (depth) *depth = (const float*)(b->h_out + b->out_dep[pair]);
    return ORB_OK;
}

int orb_batch_set_stereo(orb_batch* b, int npairs, const int* left_slots, const int* right_slots, float bf,
                         float baseline)
{
    if (!b || npairs < 0 || (npairs && (!left_slots || !right_slots))) return ORB_E_ARG;
    const int n = (int)b->ex.size();
    std::vector<StereoArgs> args(npairs);
    int waves = 0;
    for (int p = 0; p < npairs; p++) {
        const int l = left_slots[p], r = right_slots[p];
        if (l < 0 || l >= n || r < 0 || r >= n || l == r) return ORB_E_ARG;
        orb_extractor* L = b->ex[l];
        orb_extractor* R = b->ex[r];
        if (L->plan.W != R->plan.W || L->plan.H != R->plan.H || L->nlevels != R->nlevels) return ORB_E_ARG;
        if (L->plan.out_cap > 4096 || R->plan.out_cap > 4096) return ORB_E_ARG;
        StereoArgs& a = args[p];
        L->stereo_args(a, R, b->d_jobs + l, b->d_jobs + r, bf, baseline);
        a.kl = L->hjob.kps;
        a.dl = L->hjob.desc;
        a.nl = L->hjob.counters + ORB_CNT_NOUT;
        a.kr = R->hjob.kps;
        a.dr = R->hjob.desc;
        a.nr = R->hjob.counters + ORB_CNT_NOUT;
        a.cap_l = L->plan.out_cap;
        waves = std::max(waves, a.cap_l);
    }
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    b->drop_graph();
    b->drop_pinned_out();
    if (b->d_sargs) (void)hipFree(b->d_sargs);
    b->d_sargs = nullptr;
    b->pl.clear();
    b->pr.clear();
    b->h_sargs.clear();
    b->nwaves = 0;
    if (npairs == 0) return ORB_OK;
    if (hipMalloc(&b->d_sargs, npairs * sizeof(StereoArgs)) != hipSuccess) return ORB_E_OOM;
    HIP_TRY(hipMemcpy(b->d_sargs, args.data(), npairs * sizeof(StereoArgs), hipMemcpyHostToDevice));
    b->pl.assign(left_slots, left_slots + npairs);
    b->pr.assign(right_slots, right_slots + npairs);
    b->h_sargs = args;
    b->nwaves = waves;
    return ORB_OK;
}

int orb_batch_set_branches(orb_batch* b, int branches)
{
    if (!b || branches < 1 || branches > 8) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    b->drop_graph();
    b->want_branches = branches;
    return ORB_OK;
}

// The graph-branch autotuner: time mode-1 replays of the batch at each branch
// count in {1, 2, 3, 4, 6, 8}, keep the fastest (the smaller count on a tie
// within 1 %) and leave its graph built.  ms[k-1] receives the mean replay
// time at k branches, or -1 where k was not timed.
int orb_batch_autotune(orb_batch* b, int steps, int* best, float* ms)
{
    if (!b || steps <= 0 || !best) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    if (ms)
        for (int k = 0; k < 8; k++) ms[k] = -1.0f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return ORB_E_DEVICE;
    }
    int rc = ORB_OK, pick = 1;
    float pick_ms = 0.f;
    for (const int k : {1, 2, 3, 4, 6, 8}) {
        rc = orb_batch_set_branches(b, k);
        for (int w = 0; w < 2 && !rc; w++) rc = orb_batch_run(b, 1);
        if (rc) break;
        if (hipEventRecord(e0, b->ctx->stream) != hipSuccess) rc = ORB_E_DEVICE;
        for (int s = 0; s < steps && !rc; s++) rc = orb_batch_run(b, 1);
        float t = 0.f;
        if (!rc && (hipEventRecord(e1, b->ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            rc = ORB_E_DEVICE;
        if (rc) break;
        t /= (float)steps;
        if (ms) ms[k - 1] = t;
        if (k == 1 || t < 0.99f * pick_ms) {
            pick = k;
            pick_ms = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    rc = orb_batch_set_branches(b, pick);
    if (!rc) rc = orb_batch_run(b, 1);              // build the chosen graph
    if (!rc) *best = pick;
    return rc;
}

// Level kernel tile heights of every slot (levels 0 .. n-1), then the jobs
// and launch tables again.  The level tables are sized for either height, so
// no buffer moves and the slots' images stay in place.
int orb_batch_set_tile_heights(orb_batch* b, int n, const int* tile_h)
{
    if (!b || n < 0 || n > ORB_MAX_LEVELS || (n && !tile_h)) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    const int ns = (int)b->ex.size();
    std::vector<ImageJob> jobs(ns);
    for (int i = 0; i < ns; i++) {
        orb_extractor* e = b->ex[i];
        const int W = e->plan.W, H = e->plan.H;
        char* before = e->d_mem;
        int rc = orb_extractor_set_tile_heights(e, std::min(n, e->nlevels), tile_h);
        if (!rc) rc = e->ensure(W, H);
        if (rc) return rc;
        if (e->d_mem != before) return ORB_E_DEVICE;   // cannot happen: tables are sized for both heights
        jobs[i] = e->hjob;
    }
    b->drop_graph();
    if (int rc = b->refresh_pinned()) return rc;
    if (orb_extractor::build_table(b->ex.data(), ns, *b->h_rt) != ORB_OK) return ORB_E_ARG;
    HIP_TRY(hipMemcpy(b->d_jobs, jobs.data(), ns * sizeof(ImageJob), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_rt, b->h_rt, sizeof(RunTable), hipMemcpyHostToDevice));
    return ORB_OK;
}

// The tile-shape autotuner: per-level stage times (mode 0, `steps` runs)
// with 64 x 32 tiles at every level, then with 64 x 16 tiles; each level
// keeps the faster shape (64 x 16 only when it is more than 1 % faster),
// and the batch is left with that choice.  tile_h (max levels ints) gets
// it, ms32 / ms16 (optional, max levels floats) the level times.
int orb_batch_autotune_tiles(orb_batch* b, int steps, int* tile_h, float* ms32, float* ms16)
{
    if (!b || steps <= 0 || !tile_h) return ORB_E_ARG;
    const int L = b->h_rt->max_levels, ns = orb_batch_stage_count(b);
    std::vector<int> th(ORB_MAX_LEVELS, 32);
    std::vector<float> t32(ns), t16(ns);
    int rc = orb_batch_set_tile_heights(b, ORB_MAX_LEVELS, th.data());
    if (!rc) rc = orb_batch_stage_times(b, steps, t32.data());
    std::fill(th.begin(), th.end(), 16);
    if (!rc) rc = orb_batch_set_tile_heights(b, ORB_MAX_LEVELS, th.data());
    if (!rc) rc = orb_batch_stage_times(b, steps, t16.data());
    if (rc) return rc;
    for (int l = 0; l < ORB_MAX_LEVELS; l++) th[l] = l < L && t16[l] < 0.99f * t32[l] ? 16 : 32;
    rc = orb_batch_set_tile_heights(b, ORB_MAX_LEVELS, th.data());
    if (rc) return rc;
    for (int l = 0; l < L; l++) {
        tile_h[l] = th[l];
        if (ms32) ms32[l] = t32[l];
        if (ms16) ms16[l] = t16[l];
    }
    return ORB_OK;
}

int orb_batch_run(orb_batch* b, int mode)
{
    if (!b || mode < 0 || mode > 3) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    if (mode == 3) {
        // the branches launched directly on their streams, no graph
        if (b->want_branches > 1 && b->br.empty()) {
            const int rc = b->build_branches();
            if (rc) {
                b->drop_branches();
                return rc;
            }
        }
        return b->br.size() > 1 ? b->enqueue_branches() : b->enqueue(0);
    }
    if (mode != 1) return b->enqueue(mode);
    if (!b->exec) {
        // branches that direct (mode 3) launches may still be reading
        if (!b->br.empty()) HIP_TRY(hipStreamSynchronize(st));
        int rc = b->want_branches > 1 ? b->build_branches() : ORB_OK;
        if (rc) {
            b->drop_branches();
            return rc;
        }
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = b->br.size() > 1 ? b->enqueue_branches() : b->enqueue(0);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : ORB_E_DEVICE;
        }
        b->graph = g;
        if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            b->drop_graph();
            return ORB_E_DEVICE;
        }
    }
    HIP_TRY(hipGraphLaunch(b->exec, st));
    return ORB_OK;
}

int orb_batch_stage_count(const orb_batch* b)
{
    return b ? b->h_rt->max_levels + 4 : ORB_E_ARG;
}

// Per-stage device time (HIP events on the batch's stream) averaged over
// `steps` mode-0 runs: pyramid levels 0..L-1, octree, describe, stereo match,
// stereo median.  ms receives orb_batch_stage_count() floats.
int orb_batch_stage_times(orb_batch* b, int steps, float* ms)
{
    if (!b || steps <= 0 || !ms) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    const int L = b->h_rt->max_levels;
    const int nev = L + 5;  // start, L levels, octree, describe, stereo, median
    std::vector<hipEvent_t> ev(nev);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    std::vector<double> acc(nev - 1, 0.0);
    int rc = ORB_OK;
    for (int s = 0; s < steps && !rc; s++) {
        rc = b->enqueue(0, ev.data());
        if (rc) break;
        if (b->pl.empty()) {
            (void)hipEventRecord(ev[L + 3], b->ctx->stream);
            (void)hipEventRecord(ev[L + 4], b->ctx->stream);
        }
        if (hipEventSynchronize(ev[nev - 1]) != hipSuccess) rc = ORB_E_DEVICE;
        for (int i = 0; i + 1 < nev && !rc; i++) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) != hipSuccess) rc = ORB_E_DEVICE;
            acc[i] += t;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    for (int i = 0; i + 1 < nev; i++) ms[i] = (float)(acc[i] / steps);
    return ORB_OK;
}

void* orb_context_stream(orb_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int orb_batch_synchronize(orb_batch* b)
{
    if (!b) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return ORB_OK;
}

int orb_batch_results(orb_batch* b, int slot, orb_keypoint* keypoints, int capacity, uint8_t* descriptors, int* n)
{
    if (!b || slot < 0 || slot >= (int)b->ex.size() || !n) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    return b->ex[slot]->read_results(keypoints, capacity, descriptors, n);
}

int orb_batch_stereo_results(orb_batch* b, int pair, float* uRight, float* depth, int capacity, int* n)
{
    if (!b || pair < 0 || pair >= (int)b->pl.size() || !n) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(b->ctx->device));
    orb_extractor* L = b->ex[b->pl[pair]];
    int c[ORB_CNT_SLOTS];
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    HIP_TRY(hipMemcpy(c, L->hjob.counters, sizeof(c), hipMemcpyDeviceToHost));
    *n = c[ORB_CNT_NOUT];
    if (*n > capacity) return ORB_E_CAPACITY;
    if (*n > 0) {
        if (uRight) HIP_TRY(hipMemcpy(uRight, L->d_uR, (size_t)*n * 4, hipMemcpyDeviceToHost));
        if (depth) HIP_TRY(hipMemcpy(depth, L->d_depth, (size_t)*n * 4, hipMemcpyDeviceToHost));
    }
    return ORB_OK;
}

}  // extern "C"

// ============================================================= dense k=2
extern "C" {

int orb_knn2_device(orb_context* ctx, const uint8_t* d_query, int nq, const uint8_t* d_map, int64_t nm, int method,
                    float ratio, int32_t* d_idx2, int32_t* d_dist2, uint8_t* d_accept)
{
    if (!ctx || nq < 0 || nm < 0 || nm >= (1ll << 31) || (method != 0 && method != 1)) return ORB_E_ARG;
    if (nq == 0) return ORB_OK;
    if (!d_query || !d_idx2 || !d_dist2 || (nm && !d_map)) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t nchunks = (size_t)orb_knn_chunks(nm);
    int rc = ctx->grow((void**)&ctx->d_part, &ctx->part_cap, std::max<size_t>(1, (size_t)nq * nchunks) * sizeof(uint2));
    if (rc) return rc;
    HIP_TRY(orb_launch_knn2(d_query, nq, d_map, nm, method, ratio, ctx->d_part, d_idx2, d_dist2, d_accept,
                            ctx->stream, nullptr));
    return ORB_OK;
}

int orb_knn2_device_noalloc(orb_context* ctx, const uint8_t* d_query, int nq, const uint8_t* d_map, int64_t nm,
                            int method, float ratio, int32_t* d_idx2, int32_t* d_dist2, uint8_t* d_accept)
{
    if (!ctx || nq < 0 || nm < 0 || nm >= (1ll << 31) || (method != 0 && method != 1)) return ORB_E_ARG;
    if (nq == 0) return ORB_OK;
    if (!d_query || !d_idx2 || !d_dist2 || (nm && !d_map)) return ORB_E_ARG;
    size_t entries = 0;
    uint2* part = orb_knn_scratch(ctx->device, &entries);
    if (!part) return ORB_E_DEVICE;
    if ((size_t)nq * (size_t)orb_knn_chunks(nm) > entries) return ORB_E_CAPACITY;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(orb_launch_knn2(d_query, nq, d_map, nm, method, ratio, part, d_idx2, d_dist2, d_accept, ctx->stream,
                            nullptr));
    return ORB_OK;
}

int orb_knn2(orb_context* ctx, const uint8_t* query, int nq, const uint8_t* map, int64_t nm, int method, float ratio,
             int32_t* idx2, int32_t* dist2, uint8_t* accept)
{
    if (!ctx || nq < 0 || nm < 0 || (nq && (!query || !idx2 || !dist2)) || (nm && !map)) return ORB_E_ARG;
    if (nq == 0) return ORB_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ctx->grow((void**)&ctx->d_kq, &ctx->kq_cap, (size_t)nq * 32);
    if (!rc) rc = ctx->grow((void**)&ctx->d_km, &ctx->km_cap, std::max<size_t>(1, (size_t)nm * 32));
    if (rc) return rc;
    int* d_out = nullptr;
    if (hipMalloc(&d_out, (size_t)nq * 17) != hipSuccess) return ORB_E_OOM;
    hipStream_t st = ctx->stream;
    HIP_TRY(hipMemcpyAsync(ctx->d_kq, query, (size_t)nq * 32, hipMemcpyHostToDevice, st));
    if (nm) HIP_TRY(hipMemcpyAsync(ctx->d_km, map, (size_t)nm * 32, hipMemcpyHostToDevice, st));
    rc = orb_knn2_device(ctx, ctx->d_kq, nq, ctx->d_km, nm, method, ratio, d_out, d_out + 2 * (size_t)nq,
                         (uint8_t*)(d_out + 4 * (size_t)nq));
    if (!rc) {
        if (hipMemcpyAsync(idx2, d_out, (size_t)nq * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(dist2, d_out + 2 * (size_t)nq, (size_t)nq * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            (accept && hipMemcpyAsync(accept, d_out + 4 * (size_t)nq, (size_t)nq, hipMemcpyDeviceToHost, st) !=
                           hipSuccess) ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = ORB_E_DEVICE;
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_out);
    return rc;
}

// Average device time of the search kernel and of the merge kernel over
// `steps` runs.  The host query and map rows are first copied into
// context-resident device buffers; only the kernels are timed (HIP events on
// the context stream).
int orb_knn2_time(orb_context* ctx, const uint8_t* query, int nq, const uint8_t* map, int64_t nm, int method,
                  float ratio, int steps, float* ms_search, float* ms_merge)
{
    if (!ctx || steps <= 0 || nq <= 0 || !query || !map || nm <= 0 || nm >= (1ll << 31) ||
        (method != 0 && method != 1))
        return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t nchunks = (size_t)orb_knn_chunks(nm);
    int rc = ctx->grow((void**)&ctx->d_part, &ctx->part_cap, (size_t)nq * nchunks * sizeof(uint2));
    if (!rc) rc = ctx->grow((void**)&ctx->d_kq, &ctx->kq_cap, (size_t)nq * 32);
    if (!rc) rc = ctx->grow((void**)&ctx->d_km, &ctx->km_cap, (size_t)nm * 32);
    if (rc) return rc;
    int* d_out = nullptr;
    if (hipMalloc(&d_out, (size_t)nq * 17) != hipSuccess) return ORB_E_OOM;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = ORB_E_DEVICE;
    if (!rc && (hipMemcpyAsync(ctx->d_kq, query, (size_t)nq * 32, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
                hipMemcpyAsync(ctx->d_km, map, (size_t)nm * 32, hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        rc = ORB_E_DEVICE;
    double a = 0, b = 0;
    for (int s = -3; s < steps && !rc; s++) {   // s < 0: untimed warm-up (clocks ramp over the first runs)
        if (orb_launch_knn2(ctx->d_kq, nq, ctx->d_km, nm, method, ratio, ctx->d_part, d_out, d_out + 2 * (size_t)nq,
                            (uint8_t*)(d_out + 4 * (size_t)nq), ctx->stream, ev) != hipSuccess ||
            hipEventSynchronize(ev[2]) != hipSuccess) {
            rc = ORB_E_DEVICE;
            break;
        }
        float t1 = 0, t2 = 0;
        (void)hipEventElapsedTime(&t1, ev[0], ev[1]);
        (void)hipEventElapsedTime(&t2, ev[1], ev[2]);
        if (s >= 0) {
            a += t1;
            b += t2;
        }
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_out);
    if (rc) return rc;
    if (ms_search) *ms_search = (float)(a / steps);
    if (ms_merge) *ms_merge = (float)(b / steps);
    return ORB_OK;
}

}  // extern "C"

// MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cc:365-450) for
// n_points points; point p's observations are rows offsets[p] ..
// offsets[p+1]-1 of desc, in the order the reference would visit them.
static int distinctive_stage(orb_context* ctx, int n_points, const uint8_t* desc, const int32_t* offsets)
{
    const int rows = offsets[n_points];
    if (offsets[0] < 0 || rows < offsets[0]) return ORB_E_ARG;
    for (int p = 0; p < n_points; p++) {
        if (offsets[p + 1] < offsets[p]) return ORB_E_ARG;
        if (offsets[p + 1] - offsets[p] > 1024) return ORB_E_CAPACITY;
    }
    int rc = ctx->grow((void**)&ctx->d_mp, &ctx->mp_cap, (size_t)std::max(rows, 1) * 32);
    if (!rc) rc = ctx->grow((void**)&ctx->d_mpo, &ctx->mpo_cap, (size_t)(n_points + 1) * 4);
    if (!rc) rc = ctx->grow((void**)&ctx->d_mpb, &ctx->mpb_cap, (size_t)n_points * 4);
    if (!rc) rc = ctx->grow((void**)&ctx->d_mpd, &ctx->mpd_cap, (size_t)n_points * 32);
    if (rc) return rc;
    if (rows) HIP_TRY(hipMemcpyAsync(ctx->d_mp, desc, (size_t)rows * 32, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->d_mpo, offsets, (size_t)(n_points + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
    return ORB_OK;
}

extern "C" {

int orb_undistort_keypoints(orb_context* ctx, int n, const orb_keypoint* kps, const float* K, const float* D, int nD,
                            orb_keypoint* out)
{
    if (!ctx || n < 0 || !K || !D || (nD != 4 && nD != 5) || (n && (!kps || !out))) return ORB_E_ARG;
    if (n == 0) return ORB_OK;
    if (D[0] == 0.0f) {                       // Frame.cc:727: k1 == 0, no distortion
        if (out != kps) std::memmove(out, kps, (size_t)n * sizeof(orb_keypoint));
        return ORB_OK;
    }
    if (!(K[0] != 0.0f) || !(K[1] != 0.0f)) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)n * sizeof(orb_keypoint);
    int rc = ctx->grow(&ctx->d_fr, &ctx->fr_cap, 2 * bytes);
    if (rc) return rc;
    OrbKeyPoint* d_in = (OrbKeyPoint*)ctx->d_fr;
    OrbKeyPoint* d_out = d_in + n;
    const double K4[4] = {K[0], K[1], K[2], K[3]};
    double k5[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < nD; i++) k5[i] = D[i];
    hipStream_t st = ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_in, kps, bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_undistort(d_in, n, nullptr, K4, k5, d_out, st));
    HIP_TRY(hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

int orb_rgbd_depth(orb_context* ctx, int n, const orb_keypoint* kps, const orb_keypoint* kps_un, const float* depth,
                   int width, int height, size_t pitch, float bf, float* uRight, float* depth_out)
{
    if (!ctx || n < 0 || width <= 0 || height <= 0 || pitch < (size_t)width * 4 || pitch % 4 || !depth ||
        (n && (!kps || !kps_un || !uRight || !depth_out)))
        return ORB_E_ARG;
    if (n == 0) return ORB_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    // The image goes over as one linear copy of its rows at the caller's
    // pitch, which the kernel then uses (a pitched copy from pageable memory
    // would go row by row).
    const size_t kb = (size_t)n * sizeof(orb_keypoint), ib = pitch * (height - 1) + (size_t)width * 4;
    const size_t o_un = kb, o_img = 2 * kb, o_out = 2 * kb + (ib + 255) / 256 * 256;
    int rc = ctx->grow(&ctx->d_fr, &ctx->fr_cap, o_out + (size_t)n * 8);
    if (rc) return rc;
    char* m = (char*)ctx->d_fr;
    hipStream_t st = ctx->stream;
    HIP_TRY(hipMemcpyAsync(m, kps, kb, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m + o_un, kps_un, kb, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m + o_img, depth, ib, hipMemcpyHostToDevice, st));
    float* d_uR = (float*)(m + o_out);
    HIP_TRY(orb_launch_rgbd_depth((const OrbKeyPoint*)m, (const OrbKeyPoint*)(m + o_un), n, nullptr,
                                  (const float*)(m + o_img), width, height, pitch, bf, d_uR, d_uR + n, st));
    HIP_TRY(hipMemcpyAsync(uRight, d_uR, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(depth_out, d_uR + n, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

int orb_rgbd_frame(orb_extractor* ex, const uint8_t* image, int width, int height, size_t stride, const float* depth,
                   size_t depth_pitch, const float* K, const float* D, int nD, float bf, orb_keypoint* keypoints,
                   orb_keypoint* keypoints_un, int capacity, uint8_t* descriptors, float* uRight, float* depth_out,
                   int* n)
{
    if (!ex || !n || width < 0 || height < 0) return ORB_E_ARG;
    *n = 0;
    if (width == 0 || height == 0) return ORB_OK;
    if (!image || stride < (size_t)width || (depth && (depth_pitch < (size_t)width * 4 || depth_pitch % 4)) || !K ||
        !D || (nD != 4 && nD != 5) || !(K[0] != 0.0f) || !(K[1] != 0.0f))
        return ORB_E_ARG;
    orb_context* ctx = ex->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ex->ensure(width, height);
    if (rc) return rc;
    const int cap = ex->plan.out_cap;
    const size_t ib = depth ? depth_pitch * (height - 1) + (size_t)width * 4 : 0;
    const size_t o_un = (ib + 255) / 256 * 256, o_out = o_un + ((size_t)cap * sizeof(OrbKeyPoint) + 255) / 256 * 256;
    rc = ctx->grow(&ctx->d_fr, &ctx->fr_cap, o_out + (size_t)cap * 8);
    if (rc) return rc;
    rc = ex->upload_image(image, stride, hipMemcpyHostToDevice);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    char* m = (char*)ctx->d_fr;
    if (depth) HIP_TRY(hipMemcpyAsync(m, depth, ib, hipMemcpyHostToDevice, st));   // one linear copy at its pitch
    HIP_TRY(orb_launch_extract(ex->d_job, ex->d_rt, ex->h_rt, st, nullptr));
    const int* cnt = ex->hjob.counters + ORB_CNT_NOUT;
    const bool distorted = D[0] != 0.0f;                              // Frame.cc:727
    OrbKeyPoint* d_kun = distorted ? (OrbKeyPoint*)(m + o_un) : ex->hjob.kps;
    if (distorted) {
        const double K4[4] = {K[0], K[1], K[2], K[3]};
        double k5[5] = {0, 0, 0, 0, 0};
        for (int i = 0; i < nD; i++) k5[i] = D[i];
        HIP_TRY(orb_launch_undistort(ex->hjob.kps, cap, cnt, K4, k5, d_kun, st));
    }
    float* d_uR = (float*)(m + o_out);
    if (depth)
        HIP_TRY(orb_launch_rgbd_depth(ex->hjob.kps, d_kun, cap, cnt, (const float*)m, width, height, depth_pitch, bf,
                                      d_uR, d_uR + cap, st));
    rc = ex->read_results(keypoints, capacity, descriptors, n);
    if (rc) return rc;
    if (*n > 0) {
        if (keypoints_un)
            HIP_TRY(hipMemcpyAsync(keypoints_un, d_kun, (size_t)*n * sizeof(OrbKeyPoint), hipMemcpyDeviceToHost, st));
        if (depth && uRight) HIP_TRY(hipMemcpyAsync(uRight, d_uR, (size_t)*n * 4, hipMemcpyDeviceToHost, st));
        if (depth && depth_out)
            HIP_TRY(hipMemcpyAsync(depth_out, d_uR + cap, (size_t)*n * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return ORB_OK;
}

int orb_distinctive_descriptors(orb_context* ctx, int n_points, const uint8_t* desc, const int32_t* offsets,
                                int32_t* best, uint8_t* out_desc)
{
    if (!ctx || n_points < 0 || (n_points && (!offsets || !best))) return ORB_E_ARG;
    if (n_points == 0) return ORB_OK;
    if (offsets[n_points] > 0 && !desc) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = distinctive_stage(ctx, n_points, desc, offsets);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    HIP_TRY(orb_launch_distinctive(ctx->d_mp, ctx->d_mpo, n_points, ctx->d_mpb, out_desc ? ctx->d_mpd : nullptr, st));
    HIP_TRY(hipMemcpyAsync(best, ctx->d_mpb, (size_t)n_points * 4, hipMemcpyDeviceToHost, st));
    if (out_desc) HIP_TRY(hipMemcpyAsync(out_desc, ctx->d_mpd, (size_t)n_points * 32, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

int orb_distinctive_descriptors_time(orb_context* ctx, int n_points, const uint8_t* desc, const int32_t* offsets,
                                     int steps, float* ms)
{
    if (!ctx || n_points <= 0 || !desc || !offsets || steps <= 0 || !ms) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = distinctive_stage(ctx, n_points, desc, offsets);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    HIP_TRY(orb_launch_distinctive(ctx->d_mp, ctx->d_mpo, n_points, ctx->d_mpb, ctx->d_mpd, st));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    (void)hipEventRecord(e0, st);
    for (int s = 0; s < steps; s++)
        (void)orb_launch_distinctive(ctx->d_mp, ctx->d_mpo, n_points, ctx->d_mpb, ctx->d_mpd, st);
    (void)hipEventRecord(e1, st);
    const hipError_t se = hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (se != hipSuccess) return ORB_E_DEVICE;
    *ms = t / steps;
    return ORB_OK;
}

}  // extern "C"

// HBM copy microbenchmark: average milliseconds of a device-to-device copy of
// `bytes` bytes by a plain streaming kernel over `steps` runs.
extern "C" int orb_hbm_copy_time(orb_context* ctx, size_t bytes, int steps, float* ms)
{
    if (!ctx || bytes < 16 || steps <= 0 || !ms) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(ctx->device));
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return ORB_E_OOM;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return ORB_E_OOM;
    }
    int rc = ORB_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMemsetAsync(a, 1, bytes, ctx->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        rc = ORB_E_DEVICE;
    if (!rc && orb_launch_hbm_copy(a, b, bytes, ctx->stream) != hipSuccess) rc = ORB_E_DEVICE;   // warm-up
    if (!rc && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = ORB_E_DEVICE;
    for (int s = 0; s < steps && !rc; s++)
        if (orb_launch_hbm_copy(a, b, bytes, ctx->stream) != hipSuccess) rc = ORB_E_DEVICE;
    if (!rc && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = ORB_E_DEVICE;
    float t = 0.f;
    if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = ORB_E_DEVICE;
    if (!rc) *ms = t / steps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
