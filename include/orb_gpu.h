/*
 * orb_gpu.h -- C ABI of the MI355X ORB front end.
 *
 * Drop-in boundary for ORB-SLAM2's data-parallel hot path.  Each entry point
 * names the reference interface it replaces (ORB-SLAM2 file:line).  Plain
 * pointers and sizes only; every call returns an int status (ORB_OK or a
 * negative ORB_E* code) and never throws.
 *
 * Ownership: the caller owns host buffers; device buffers belong to the
 * context or extractor handle.  A handle is not thread-safe; distinct handles
 * may be used concurrently from different threads.
 */
#ifndef ORB_GPU_H
#define ORB_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORB_OK 0
#define ORB_E_ARG (-1)       /* bad argument (null, size, parameter range) */
#define ORB_E_CAPACITY (-2)  /* caller's output buffer too small */
#define ORB_E_DEVICE (-3)    /* HIP runtime error */
#define ORB_E_OOM (-4)       /* device allocation failed */
#define ORB_E_OVERFLOW (-5)  /* an internal bound was exceeded (see orb_extractor_capacity) */
#define ORB_E_ADMISSION (-6) /* serving: request refused by the memory budget */

/* Same layout as cv::KeyPoint: pt.x, pt.y, size, angle, response, octave,
 * class_id (28 bytes). */
typedef struct orb_keypoint {
    float x, y, size, angle, response;
    int octave, class_id;
} orb_keypoint;

typedef struct orb_context orb_context;
typedef struct orb_extractor orb_extractor;

/* Device context: binds a HIP device and owns the stream all work runs on.
 * No reference counterpart (the reference is CPU-only).  Every extractor,
 * batch, server, vocabulary and rectifier made on a context holds a
 * reference to it: orb_context_destroy releases the caller's, and the context
 * is freed with the last of them, so handles may be destroyed in any order. */
int orb_context_create(int device, orb_context** out);
void orb_context_destroy(orb_context* ctx);
int orb_context_synchronize(orb_context* ctx);
/* The context's hipStream_t (as void*), for callers that time or order work
 * around the library's launches. */
void* orb_context_stream(orb_context* ctx);

/* ORBextractor::ORBextractor(nfeatures, scaleFactor, nlevels, iniThFAST,
 * minThFAST)  (include/ORBextractor.h:93, src/ORBextractor.cc:469-571).
 * nlevels <= 16, scaleFactor > 1. */
int orb_extractor_create(orb_context* ctx, int nfeatures, float scaleFactor, int nlevels, int iniThFAST,
                         int minThFAST, orb_extractor** out);
void orb_extractor_destroy(orb_extractor* ex);

/* How resize and the 8U Gaussian round, i.e. which OpenCV build the results
 * match.  No counterpart in the reference, whose rounding is its OpenCV
 * library's: mode 1 (the default) = the SSE2 kernels every x86 build of
 * OpenCV 2.4/3.x runs (resize's VResizeLinearVec_32s8u, the Gaussian's
 * SymmColumnVec_32s8u), mode 0 = OpenCV's generic scalar code (builds
 * without SSE2).  Takes effect at the next extraction. */
int orb_extractor_set_opencv_rounding(orb_extractor* ex, int mode);
/* GetLevels / GetScaleFactor (ORBextractor.h:119-124). */
int orb_extractor_levels(const orb_extractor* ex);
float orb_extractor_scale_factor(const orb_extractor* ex);
/* GetScaleFactors / GetInverseScaleFactors / GetScaleSigmaSquares /
 * GetInverseScaleSigmaSquares (ORBextractor.h:127-159); any pointer may be
 * NULL; each non-NULL one receives nlevels floats. */
int orb_extractor_scales(const orb_extractor* ex, float* scale, float* inv_scale, float* sigma2,
                         float* inv_sigma2);
/* mnFeaturesPerLevel (ORBextractor.cc:509-526): nlevels ints. */
int orb_extractor_quotas(const orb_extractor* ex, int* quota);
/* Upper bound on keypoints per call: sum over levels of max(quota+3, 4*roots),
 * roots from the aspect ratio of width x height. */
int orb_extractor_capacity(const orb_extractor* ex, int width, int height);

/* ORBextractor::operator()(image, mask, keypoints, descriptors)
 * (include/ORBextractor.h:111, src/ORBextractor.cc:1533-1649).
 * image: width x height CV_8UC1 host pixels with row stride `stride` bytes.
 * On return *n keypoints (level-0 coordinates, octave order then quadtree
 * order) and n x 32 descriptor bytes.  An empty image (width or height 0)
 * returns ORB_OK with *n = 0, as the reference returns silently.  The mask
 * argument of the reference is ignored by the reference and is not taken. */
int orb_extract(orb_extractor* ex, const uint8_t* image, int width, int height, size_t stride,
                orb_keypoint* keypoints, int capacity, uint8_t* descriptors, int* n);

/* Same, with the image already in device memory (row pitch in bytes).  The
 * results stay on the device until read with orb_extractor_results; the call
 * returns once the work is enqueued. */
int orb_extract_device(orb_extractor* ex, const uint8_t* d_image, int width, int height, size_t pitch);
int orb_extractor_results(orb_extractor* ex, orb_keypoint* keypoints, int capacity, uint8_t* descriptors,
                          int* n);

/* Colour input, converted on the device in front of the pyramid exactly as
 * cvtColor does for 8-bit images (src/Tracking.cc:250-276): channels 3 or 4,
 * rgb = 1 for RGB/RGBA (CV_RGB2GRAY, CV_RGBA2GRAY), 0 for BGR/BGRA;
 * gray = (R*4899 + G*9617 + B*1868 + 8192) >> 14.  Otherwise as orb_extract. */
int orb_extract_color(orb_extractor* ex, const uint8_t* image, int width, int height, size_t stride, int channels,
                      int rgb, orb_keypoint* keypoints, int capacity, uint8_t* descriptors, int* n);
/* The conversion alone, into a caller's gray buffer. */
int orb_cvt_color_gray(orb_context* ctx, const uint8_t* image, int width, int height, size_t stride, int channels,
                       int rgb, uint8_t* gray, size_t gray_stride);

/* Rectification in front of the pyramid (Examples/Stereo/stereo_euroc.cc
 * :124-188).  orb_rectify_create computes cv::initUndistortRectifyMap(K, D,
 * R, P, (width, height), CV_32FC1) on the host in double (K, R, P row-major
 * 3x3, R may be NULL for identity; D has 0, 4, 5, 8 or 12 coefficients in
 * OpenCV's order k1 k2 p1 p2 k3 k4 k5 k6 s1 s2 s3 s4) and keeps the maps in
 * HBM; orb_rectify_from_maps takes maps the caller already has.  orb_remap is
 * cv::remap(INTER_LINEAR) with the default zero border; orb_extract_rectified
 * remaps straight into the extractor's input and extracts (the output size is
 * the map's). */
typedef struct orb_rectify orb_rectify;
int orb_rectify_create(orb_context* ctx, const double* K, const double* D, int nD, const double* R, const double* P,
                       int width, int height, orb_rectify** out);
int orb_rectify_from_maps(orb_context* ctx, const float* mapx, const float* mapy, int width, int height,
                          orb_rectify** out);
void orb_rectify_destroy(orb_rectify* r);
int orb_rectify_maps(orb_rectify* r, float* mapx, float* mapy);
int orb_remap(orb_rectify* r, const uint8_t* src, int src_width, int src_height, size_t src_stride, uint8_t* dst,
              size_t dst_stride);
int orb_extract_rectified(orb_extractor* ex, orb_rectify* r, const uint8_t* src, int src_width, int src_height,
                          size_t src_stride, orb_keypoint* keypoints, int capacity, uint8_t* descriptors, int* n);

/* mvImagePyramid[level] (ORBextractor.h:162) of the last call, copied to a
 * host buffer (dst may be NULL to query the size). */
int orb_extractor_pyramid_level(orb_extractor* ex, int level, uint8_t* dst, size_t dst_stride, int* width,
                                int* height);

/* ORBmatcher::DescriptorDistance(a, b) (include/ORBmatcher.h:65,
 * src/ORBmatcher.cc:2083-2103): Hamming distance of two 32-byte rows. */
int orb_descriptor_distance(const uint8_t* a, const uint8_t* b);

/* Brute-force matching of nq query rows against nt train rows with the
 * best/second-best rule of ORBmatcher::SearchForInitialization
 * (src/ORBmatcher.cc:614-647) and no spatial window: idx[i] = best train index
 * if best <= th_low and best < (float)second * ratio, else -1; dist[i] = best. */
int orb_match_bf(orb_context* ctx, const uint8_t* query, int nq, const uint8_t* train, int nt, int th_low,
                 float ratio, int32_t* idx, int32_t* dist);

/* Dense k=2 search: for each of nq query rows the two map rows nearest in
 * Hamming distance, ordered by (distance, index) -- the order of the
 * best/second-best update of ORBmatcher (src/ORBmatcher.cc:614-647, distance
 * of ORBmatcher::DescriptorDistance) -- over all nm map rows.
 * idx2/dist2: 2 ints per query (-1 / INT_MAX when fewer than two rows);
 * accept[i] = best < (float)second * ratio (may be NULL).
 * method 0: XOR + popcount kernel; 1: int8 +-1 MFMA kernel (dot = 256 - 2*d).
 * Both give identical results.  nm < 2^31. */
int orb_knn2(orb_context* ctx, const uint8_t* query, int nq, const uint8_t* map, int64_t nm, int method, float ratio,
             int32_t* idx2, int32_t* dist2, uint8_t* accept);
/* Same on device-resident buffers; returns once enqueued on the context stream. */
int orb_knn2_device(orb_context* ctx, const uint8_t* d_query, int nq, const uint8_t* d_map, int64_t nm, int method,
                    float ratio, int32_t* d_idx2, int32_t* d_dist2, uint8_t* d_accept);
/* Allocation-free form of orb_knn2_device.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-chunk partial results live in a fixed 4 MiB array per GPU (2^19 keys),
 * so nq * ceil(nm / 8192) <= 524288 (ORB_E_CAPACITY otherwise: e.g. 2000
 * queries against up to 2.1 M map rows).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other contexts or streams) race. */
int orb_knn2_device_noalloc(orb_context* ctx, const uint8_t* d_query, int nq, const uint8_t* d_map, int64_t nm,
                            int method, float ratio, int32_t* d_idx2, int32_t* d_dist2, uint8_t* d_accept);
/* Copies query and map rows into HBM, then returns the average device
 * milliseconds of the search and merge kernels over `steps` runs after three
 * untimed ones (HIP events on the context stream). */
int orb_knn2_time(orb_context* ctx, const uint8_t* query, int nq, const uint8_t* map, int64_t nm, int method,
                  float ratio, int steps, float* ms_search, float* ms_merge);

/* Frame::ComputeStereoMatches() (include/Frame.h:249, src/Frame.cc:831-1082)
 * on keypoints/descriptors of a left and a right image, reading the two
 * extractors' pyramids from their last calls.  bf = Camera.bf, b = bf/fx.
 * uRight/depth receive nl floats each (-1 where unmatched).  Both extractors
 * must belong to the same context (ORB_E_ARG otherwise). */
int orb_stereo_match(orb_extractor* left, orb_extractor* right, const orb_keypoint* kl, const uint8_t* dl,
                     int nl, const orb_keypoint* kr, const uint8_t* dr, int nr, float bf, float b, float* uRight,
                     float* depth);

/* The stereo Frame constructor's hot path (src/Frame.cc:146-169): extract
 * both images in the same launches, then ComputeStereoMatches on the device
 * with no host round trip. */
int orb_stereo_frame(orb_extractor* left, orb_extractor* right, const uint8_t* im_left, const uint8_t* im_right,
                     int width, int height, size_t stride, float bf, float b, orb_keypoint* kl, uint8_t* dl,
                     int* nl, orb_keypoint* kr, uint8_t* dr, int* nr, int capacity, float* uRight, float* depth);

/* ---------------------------------------------------------------- batches
 * A fixed set of image slots (each with its own size and nfeatures) extracted
 * together.  No reference counterpart: the reference extracts one image per
 * call; this is the batched form of ORBextractor::operator() for many frames
 * and cameras.  Images are written into the slots' device buffers (resident
 * in HBM), then orb_batch_run enqueues the whole pipeline:
 *   mode 0  one launch per stage covering every slot
 *   mode 1  the mode-0 launches captured once into a hipGraph and replayed;
 *           with orb_batch_set_branches(b, k > 1) the slots are split into k
 *           groups (whole stereo pairs) whose launches form k independent
 *           graph branches, so one group's small, short launches overlap
 *           another's large ones
 *   mode 2  per-image launches of the same kernels (the unbatched baseline)
 *   mode 3  the branches of mode 1 launched directly on their own streams,
 *           without a graph (one stream per branch, joined into the
 *           context's stream)
 * Stereo pairs declared with orb_batch_set_stereo are matched after
 * extraction (Frame::ComputeStereoMatches, src/Frame.cc:831-1082). */
typedef struct orb_batch orb_batch;
int orb_batch_create(orb_context* ctx, int n, const int* widths, const int* heights, const int* nfeatures,
                     float scaleFactor, int nlevels, int iniThFAST, int minThFAST, orb_batch** out);
void orb_batch_destroy(orb_batch* b);
int orb_batch_size(const orb_batch* b);
/* Device pointer, pitch and size of a slot's level-0 image buffer. */
int orb_batch_image(orb_batch* b, int slot, uint8_t** d_image, size_t* pitch, int* width, int* height);
int orb_batch_upload(orb_batch* b, int slot, const uint8_t* image, size_t stride);
int orb_batch_set_stereo(orb_batch* b, int npairs, const int* left_slots, const int* right_slots, float bf,
                         float baseline);
int orb_batch_run(orb_batch* b, int mode);
int orb_batch_synchronize(orb_batch* b);
/* Number of graph branches for mode 1 (1..8, default 1).  Slots are cut into
 * contiguous groups of near-equal size that keep every stereo pair whole; if
 * no such cut exists the graph has fewer branches.  Results are identical. */
int orb_batch_set_branches(orb_batch* b, int branches);
/* Autotuner of the graph's branch count: times `steps` mode-1 replays at 1, 2,
 * 3, 4, 6 and 8 branches (whatever the slots' current images), keeps the
 * fastest (the smaller count within 1 %), builds its graph and returns it in
 * *best.  ms (8 floats, may be NULL) receives the mean replay time in ms at k
 * branches in ms[k-1], or -1 where k was not timed. */
int orb_batch_autotune(orb_batch* b, int steps, int* best, float* ms);
/* Tile shape of the level kernel, per pyramid level: 64 x 32 (the default)
 * or 64 x 16 tiles (tile_h[l] = 32 or 16, levels 0 .. n-1).  No counterpart
 * in the reference; results are identical either way.  Smaller tiles halve
 * each workgroup's serial work, which shortens the small levels, whose few
 * tiles leave the GPU partly idle; larger ones share their halo over more
 * pixels.  On a batch it applies to every slot, whose images stay in place. */
int orb_extractor_set_tile_heights(orb_extractor* ex, int n, const int* tile_h);
int orb_batch_set_tile_heights(orb_batch* b, int n, const int* tile_h);
/* Autotuner of the tile shapes: times the pyramid stages (mode 0, `steps`
 * runs) with 64 x 32 and with 64 x 16 tiles at every level, keeps per level
 * the faster one (64 x 16 only when more than 1 % faster) and leaves the
 * batch with that choice.  tile_h receives it (one int per level), ms32 /
 * ms16 (optional) the level times in ms. */
int orb_batch_autotune_tiles(orb_batch* b, int steps, int* tile_h, float* ms32, float* ms16);
/* Stage timing with HIP events on the batch's stream, averaged over `steps`
 * mode-0 runs.  ms receives orb_batch_stage_count() floats: pyramid levels
 * 0..L-1 (pyramid, blur, FAST with per-cell suppression), quadtree, describe,
 * stereo match, stereo median. */
int orb_batch_stage_count(const orb_batch* b);
int orb_batch_stage_times(orb_batch* b, int steps, float* ms);
int orb_batch_results(orb_batch* b, int slot, orb_keypoint* keypoints, int capacity, uint8_t* descriptors, int* n);
int orb_batch_stereo_results(orb_batch* b, int pair, float* uRight, float* depth, int capacity, int* n);
/* Whole-batch host transfers (no reference counterpart; the per-slot calls
 * above wait once per slot).  orb_batch_upload_images enqueues the copies of
 * n host images into the given slots on the batch's stream and returns: the
 * next orb_batch_run is ordered after them, and the host buffers may be
 * reused on return.  orb_batch_read_all waits for the batch, then copies
 * every slot's keypoints and descriptors (slot s at keypoints + s *
 * capacity, descriptors + s * capacity * 32, counts[s] of them) and, for
 * stereo pairs, every pair's uRight / depth (pair p at uRight + p * capacity;
 * either may be NULL), with one wait for the counts and one for the data. */
int orb_batch_upload_images(orb_batch* b, int n, const int* slots, const uint8_t* const* images,
                            const size_t* strides);
int orb_batch_read_all(orb_batch* b, orb_keypoint* keypoints, uint8_t* descriptors, int capacity, int* counts,
                       float* uRight, float* depth);
/* Pinned whole-batch transfers (no reference counterpart), for frames that a
 * capture or decode step writes straight into pinned host memory and for
 * pipelines that keep several batches in flight: nothing here waits.
 *   orb_batch_pinned_images: one pinned host buffer per slot (rows of
 *     pitches[s] >= width bytes), allocated by the first call and owned by
 *     the batch; write frame s at images[s].
 *   orb_batch_upload_pinned: enqueue one host-to-device copy of all of them
 *     and one launch that moves each into its slot's image; the next
 *     orb_batch_run is ordered after it.  The host buffers may be rewritten
 *     once the batch's stream has passed it (orb_batch_synchronize, or a
 *     later orb_batch_pinned_results).
 *   orb_batch_read_async: enqueue one launch that gathers every slot's
 *     counters, keypoints and descriptors (valid entries only) and every
 *     pair's uRight / depth into one device buffer, and one device-to-host
 *     copy of it into pinned memory the batch owns.
 *   orb_batch_pinned_results / orb_batch_pinned_stereo: wait for the last
 *     orb_batch_read_async, then point into its pinned copy (n entries; valid
 *     until the next read, stereo, tile-height change or destroy).
 *     ORB_E_ARG if no read was enqueued since the last such change,
 *     ORB_E_OVERFLOW as orb_batch_results. */
int orb_batch_pinned_images(orb_batch* b, uint8_t** images, size_t* pitches);
int orb_batch_upload_pinned(orb_batch* b);
int orb_batch_read_async(orb_batch* b);
int orb_batch_pinned_results(orb_batch* b, int slot, const orb_keypoint** keypoints, const uint8_t** descriptors,
                             int* n);
int orb_batch_pinned_stereo(orb_batch* b, int pair, const float** uRight, const float** depth, int* n);

/* ---------------------------------------------------------------- serving
 * Continuous-batching request loop for many camera streams.  No reference
 * counterpart: it serves the per-frame hot path (ORB extraction as in the
 * Frame constructors, src/Frame.cc:121-170, plus a brute-force Hamming match
 * against the stream's previous frame, best/second rule of
 * src/ORBmatcher.cc:614-647) for many cameras at once.  A server thread admits
 * pending frames, oldest first, into extractor slots while the HBM budget
 * allows (one frame per stream in flight, so each stream stays ordered),
 * launches them as one batch on a free HIP stream (up to max_inflight batches
 * overlap), keeps every stream's last descriptors resident in HBM, and stamps
 * latencies with orb_server_clock() (CLOCK_MONOTONIC seconds).
 * A frame whose size can never fit the budget completes with
 * status ORB_E_ADMISSION; submit refuses (ORB_E_ADMISSION) beyond 4096
 * pending frames. */
typedef struct orb_server orb_server;
typedef struct orb_serve_info {
    uint64_t ticket;
    int stream, status, n_keypoints, n_matches;
    double t_arrival, t_admit, t_done;
} orb_serve_info;
int orb_server_create(orb_context* ctx, int nstreams, const int* widths, const int* heights, const int* nfeatures,
                      float scaleFactor, int nlevels, int iniThFAST, int minThFAST, int th_low, float ratio,
                      size_t hbm_budget, int max_inflight, int max_batch, orb_server** out);
void orb_server_destroy(orb_server* s);
double orb_server_clock(void);
int orb_server_submit(orb_server* s, int stream, const uint8_t* image, size_t stride, double t_arrival,
                      uint64_t* ticket);
/* Completed frames, oldest completion first. */
int orb_server_poll(orb_server* s, orb_serve_info* out, int max, int* n);
/* Results of a polled frame (then released): keypoints, descriptors and the
 * match of each keypoint into the stream's previous frame (-1 = none). */
int orb_server_fetch(orb_server* s, uint64_t ticket, orb_keypoint* keypoints, int capacity, uint8_t* descriptors,
                     int32_t* match_idx, int32_t* match_dist, int* n);
int orb_server_release(orb_server* s, uint64_t ticket);
int orb_server_drain(orb_server* s, double timeout_s);
int orb_server_stats(orb_server* s, int64_t* batches, int64_t* requests, size_t* hbm_used, int* slots);
/* Load generator (no reference counterpart): an open-loop replay from native
 * threads, so that the server is measured without a scripting-language
 * driver.  Streams 0 .. nstreams-1 each get a submitting thread; stream s
 * replays frames[s] (nframes[s] images of its size, rows contiguous) forwards
 * then backwards, at exponential inter-arrivals of mean 1/fps (generator
 * seeded with seed + s) for `duration` seconds, and the calling thread polls
 * and releases until every submitted frame has completed.  Latency is
 * completion minus scheduled arrival; frames_per_s counts completions from
 * the first arrival to the last completion. */
typedef struct orb_replay_stats {
    int64_t requests, refused, errors;
    double frames_per_s, p50_ms, p99_ms, max_ms;
} orb_replay_stats;
int orb_server_replay(orb_server* s, int nstreams, const uint8_t* const* frames, const int* nframes, double fps,
                      double duration, uint64_t seed, orb_replay_stats* out);

/* ---------------------------------------------------------------- vocabulary
 * DBoW2 bag of words (Thirdparty/DBoW2/DBoW2/TemplatedVocabulary.h), as
 * Frame::ComputeBoW / KeyFrame::ComputeBoW call it (src/Frame.cc:704-718,
 * levelsup 4).  The tree lives in HBM, renumbered breadth-first; node ids in
 * the outputs are the file's. */
typedef struct orb_vocabulary orb_vocabulary;
/* TemplatedVocabulary::loadFromTextFile (TemplatedVocabulary.h:1368): the
 * ORBvoc.txt format.  Blank lines are skipped. */
int orb_vocabulary_load_text(orb_context* ctx, const char* path, orb_vocabulary** out);
/* The same tree from arrays in file order: n_nodes entries, entry 0 the root
 * (its fields ignored), parent[i] < i, is_word[i] > 0 for words, 32-byte
 * descriptors, weights.  scoring 0..5 (L1_NORM .. DOT_PRODUCT), weighting
 * 0..3 (TF_IDF, TF, IDF, BINARY), as DBoW2's enums. */
int orb_vocabulary_create(orb_context* ctx, int k, int L, int scoring, int weighting, int n_nodes,
                          const int32_t* parent, const uint8_t* is_word, const uint8_t* desc, const double* weight,
                          orb_vocabulary** out);
void orb_vocabulary_destroy(orb_vocabulary* v);
int orb_vocabulary_info(const orb_vocabulary* v, int* k, int* L, int* scoring, int* weighting, int* n_nodes,
                        int* n_words);
/* TemplatedVocabulary::transform(feature, word_id, weight, nid, levelsup)
 * (TemplatedVocabulary.h:1240-1290) for n descriptors: word id, word weight
 * and the node at depth L - levelsup (node_id may be NULL). */
int orb_vocabulary_transform(orb_vocabulary* v, const uint8_t* desc, int n, int levelsup, int32_t* word_id,
                             double* weight, int32_t* node_id);
/* TemplatedVocabulary::transform(features, BowVector, FeatureVector, levelsup)
 * (TemplatedVocabulary.h:1136-1206) for n_images images at once.  Image i is
 * rows offsets[i] .. offsets[i+1]-1 of desc (at most 4096 rows); its outputs
 * are written from the same offset: the BowVector as n_bow[i] (word id,
 * value) pairs in word order, the FeatureVector as n_fv[i] (node id, feature
 * index) pairs in (node, feature) order. */
int orb_vocabulary_bow(orb_vocabulary* v, int n_images, const uint8_t* desc, const int32_t* offsets, int levelsup,
                       int32_t* bow_word, double* bow_value, int32_t* n_bow, int32_t* fv_node, int32_t* fv_feature,
                       int32_t* n_fv);
/* Measurement aid, no reference counterpart: average milliseconds of the
 * device work of orb_vocabulary_bow with the descriptors resident in HBM. */
int orb_vocabulary_bow_time(orb_vocabulary* v, int n_images, const uint8_t* desc, const int32_t* offsets,
                            int levelsup, int steps, float* ms);

/* Windowed searches (SearchForInitialization, SearchByProjection) resolve
 * the reference's in-order passes in parallel rounds, up to `rounds` (0..48,
 * default 48) before running the in-order pass for what has not settled;
 * results are identical either way.  No reference counterpart: a tuning and
 * test aid.  orb_context_last_rounds returns the rounds the last windowed
 * search on this context ran and sets *sequential (may be NULL) to 1 if it
 * fell back to the in-order pass for some pair or frame. */
int orb_context_set_round_limit(orb_context* ctx, int rounds);
int orb_context_last_rounds(const orb_context* ctx, int* sequential);

/* ---------------------------------------------------------------- windowed matching
 * ORBmatcher::SearchForInitialization (src/ORBmatcher.cc:573-717) over the
 * Frame keypoint grid (64 x 48 cells, src/Frame.cc:399-423, 590-700), for
 * n_pairs frame pairs at once.  Pair p's F1 keypoints (mvKeysUn) and
 * descriptors are rows off1[p] .. off1[p+1]-1, F2's rows off2[p] ..
 * off2[p+1]-1 (at most 4096 each); bounds holds mnMinX, mnMaxX, mnMinY,
 * mnMaxY per pair; prev_matched (x, y per F1 keypoint) is vbPrevMatched,
 * updated in place for the matched keypoints.  The initializer calls it with
 * window 100, ratio 0.9 and the orientation check (Tracking.cc:915-927).
 * matches12 receives vnMatches12 per F1 keypoint, nmatches the count per
 * pair.  The reference's in-order pass over the F1 keypoints runs as parallel
 * rounds that converge to its result; pairs still changing after the
 * context's round limit get the in-order pass itself. */
int orb_search_for_initialization(orb_context* ctx, int n_pairs, const orb_keypoint* kp1, const uint8_t* desc1,
                                  const int32_t* off1, const orb_keypoint* kp2, const uint8_t* desc2,
                                  const int32_t* off2, const float* bounds, float* prev_matched, int window_size,
                                  float nn_ratio, int check_orientation, int32_t* matches12, int32_t* nmatches);
/* Measurement aid, no reference counterpart: the call above, then `steps`
 * more runs of its device work with the inputs resident in HBM (host reads
 * of the candidate counts and change flags included, vbPrevMatched restored
 * before each), their mean wall time in *ms.  The outputs are the call's. */
int orb_search_for_initialization_time(orb_context* ctx, int n_pairs, const orb_keypoint* kp1, const uint8_t* desc1,
                                       const int32_t* off1, const orb_keypoint* kp2, const uint8_t* desc2,
                                       const int32_t* off2, const float* bounds, float* prev_matched, int window_size,
                                       float nn_ratio, int check_orientation, int32_t* matches12, int32_t* nmatches,
                                       int steps, float* ms);

/* ORBmatcher::SearchByProjection, both per-frame overloads, for n_frames
 * frames at once.  Frame f's keypoints (mvKeysUn, at most 4096),
 * descriptors, mvuRight and occupied[i] = (mvpMapPoints[i] &&
 * mvpMapPoints[i]->Observations() > 0) are rows key_off[f] .. key_off[f+1]-1;
 * its points are rows point_off[f] .. point_off[f+1]-1 of points /
 * point_desc (MapPoint::GetDescriptor), in the order the reference visits
 * them; bounds holds mnMinX, mnMaxX, mnMinY, mnMaxY per frame and
 * scale_factors the frame's mvScaleFactors.  Outputs: point_match[p] = the
 * keypoint point p was assigned (-1 none); key_point[i] = -1 untouched, >= 0
 * the point (index within its frame) now in mvpMapPoints[i], -2 set to NULL;
 * nmatches per frame.  The caller applies key_point to mvpMapPoints.  A later
 * point skips keypoints an earlier point with observations took, exactly as
 * the reference's in-order pass. */

/* One entry of vpMapPoints as Frame::isInFrustum (src/Frame.cc:491-578) left
 * it.  flags: 1 mbTrackInView, 2 isBad(), 4 Observations() > 0. */
typedef struct orb_proj_point {
    float proj_x, proj_y, proj_xr;   /* mTrackProjX, mTrackProjY, mTrackProjXR */
    float view_cos;                  /* mTrackViewCos */
    int level;                       /* mnTrackScaleLevel */
    int flags;
} orb_proj_point;

/* ORBmatcher::SearchByProjection(Frame &F, const vector<MapPoint*>
 * &vpMapPoints, const float th) (include/ORBmatcher.h:82, src/ORBmatcher.cc:72-183,
 * RadiusByViewingCos :186-202); nn_ratio is the matcher's mfNNratio (0.8 in
 * Tracking::SearchLocalPoints). */
int orb_search_by_projection(orb_context* ctx, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                             const float* uright, const uint8_t* occupied, const int32_t* key_off,
                             const float* bounds, const float* scale_factors, int nlevels,
                             const orb_proj_point* points, const uint8_t* point_desc, const int32_t* point_off,
                             float th, float nn_ratio, int32_t* point_match, int32_t* key_point, int32_t* nmatches);

/* LastFrame.mvpMapPoints[i] in index order: its position in the current
 * camera frame (Rcw * x3Dw + tcw, computed by the caller), the last frame's
 * mvKeysUn[i].angle and mvKeys[i].octave.  flags: 1 mvpMapPoints[i] &&
 * !mvbOutlier[i], 4 Observations() > 0. */
typedef struct orb_last_point {
    float xc, yc, zc;
    float angle;
    int octave;
    int flags;
} orb_last_point;

/* ORBmatcher::SearchByProjection(Frame &CurrentFrame, const Frame &LastFrame,
 * const float th, const bool bMono) (include/ORBmatcher.h:102,
 * src/ORBmatcher.cc:1710-1889).  camera holds per frame fx, fy, cx, cy, mbf,
 * mb and tlc.z (the z of the last-to-current translation Rlw * twc + tlw,
 * computed by the caller); check_orientation is mbCheckOrientation. */
int orb_search_by_projection_last(orb_context* ctx, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                                  const float* uright, const uint8_t* occupied, const int32_t* key_off,
                                  const float* bounds, const float* scale_factors, int nlevels, const float* camera,
                                  int mono, const orb_last_point* points, const uint8_t* point_desc,
                                  const int32_t* point_off, float th, int check_orientation, int32_t* point_match,
                                  int32_t* key_point, int32_t* nmatches);
/* Measurement aids, no reference counterpart: the calls above, then `steps`
 * more runs of their device work with the inputs resident in HBM (host reads
 * of the candidate counts and change flags included), their mean wall time in
 * *ms.  The outputs are the call's. */
int orb_search_by_projection_time(orb_context* ctx, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                                  const float* uright, const uint8_t* occupied, const int32_t* key_off,
                                  const float* bounds, const float* scale_factors, int nlevels,
                                  const orb_proj_point* points, const uint8_t* point_desc, const int32_t* point_off,
                                  float th, float nn_ratio, int32_t* point_match, int32_t* key_point,
                                  int32_t* nmatches, int steps, float* ms);
int orb_search_by_projection_last_time(orb_context* ctx, int n_frames, const orb_keypoint* kp, const uint8_t* desc,
                                       const float* uright, const uint8_t* occupied, const int32_t* key_off,
                                       const float* bounds, const float* scale_factors, int nlevels,
                                       const float* camera, int mono, const orb_last_point* points,
                                       const uint8_t* point_desc, const int32_t* point_off, float th,
                                       int check_orientation, int32_t* point_match, int32_t* key_point,
                                       int32_t* nmatches, int steps, float* ms);

/* ---------------------------------------------------------------- Frame keypoint ops
 * Frame::UndistortKeyPoints() (include/Frame.h, src/Frame.cc:724-776): the
 * keypoints' cv::undistortPoints(pts, pts, mK, mDistCoef, Mat(), mK).
 * K = fx, fy, cx, cy; D = nD (4 or 5) coefficients k1 k2 p1 p2 [k3], as
 * mDistCoef holds them.  out receives the keypoints with x, y undistorted
 * (every other field copied; out may be kps).  k1 == 0 copies, as the
 * reference does (Frame.cc:727). */
int orb_undistort_keypoints(orb_context* ctx, int n, const orb_keypoint* kps, const float* K, const float* D, int nD,
                            orb_keypoint* out);
/* Frame::ComputeStereoFromRGBD() (src/Frame.cc:1085-1123): per keypoint,
 * d = depth at row (int)y, column (int)x of the keypoint as extracted (the
 * float depth image of width x height, `pitch` bytes a row, as Tracking has
 * scaled it); d > 0 gives depth_out = d and uRight = kps_un.x - bf / d, else
 * both -1.  A keypoint outside the image (read out of bounds by the
 * reference) gets -1. */
int orb_rgbd_depth(orb_context* ctx, int n, const orb_keypoint* kps, const orb_keypoint* kps_un, const float* depth,
                   int width, int height, size_t pitch, float bf, float* uRight, float* depth_out);
/* The RGB-D Frame constructor's hot path in one call (src/Frame.cc:221-263:
 * ExtractORB, UndistortKeyPoints, ComputeStereoFromRGBD): the extraction as
 * orb_extract, then the undistortion and the depth lookup on the keypoints
 * still on the device, one wait at the end.  keypoints / descriptors as
 * orb_extract; keypoints_un, uRight and depth_out (each optional) receive
 * *n entries.  depth: width x height floats, depth_pitch bytes a row; NULL
 * for the monocular constructor (src/Frame.cc:174-216), which stops after
 * the undistortion (uRight and depth_out are then not written). */
int orb_rgbd_frame(orb_extractor* ex, const uint8_t* image, int width, int height, size_t stride, const float* depth,
                   size_t depth_pitch, const float* K, const float* D, int nD, float bf, orb_keypoint* keypoints,
                   orb_keypoint* keypoints_un, int capacity, uint8_t* descriptors, float* uRight, float* depth_out,
                   int* n);

/* ---------------------------------------------------------------- map points
 * MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cc:365-450) for
 * n_points points at once.  Point p's observation descriptors are rows
 * offsets[p] .. offsets[p+1]-1 of desc (at most 1024), in the order the
 * reference visits its observations.  best[p] receives the index, within
 * the point's rows, of the descriptor with the least median distance to the
 * others (the first on ties), or -1 for a point without rows; out_desc
 * (optional, n_points x 32) receives that descriptor (zeros for -1). */
int orb_distinctive_descriptors(orb_context* ctx, int n_points, const uint8_t* desc, const int32_t* offsets,
                                int32_t* best, uint8_t* out_desc);
/* Measurement aid, no reference counterpart: average milliseconds of the
 * device work of orb_distinctive_descriptors with the rows resident in HBM. */
int orb_distinctive_descriptors_time(orb_context* ctx, int n_points, const uint8_t* desc, const int32_t* offsets,
                                     int steps, float* ms);

/* Measurement aid, no reference counterpart: average milliseconds of a
 * device-to-device copy of `bytes` bytes by a plain streaming kernel (HBM
 * bandwidth check and hardware-counter calibration). */
int orb_hbm_copy_time(orb_context* ctx, size_t bytes, int steps, float* ms);

/* Debug accessors used by the parity tests.
 * kind 0: pyramid level, 1: its 7x7 blur. */
int orb_extractor_debug_image(orb_extractor* ex, int level, int kind, uint8_t* dst, size_t dst_stride);
/* FAST candidates of a level as (x, y, response) triples relative to the
 * level's 16-px border, sorted in the reference's vToDistributeKeys order. */
int orb_extractor_debug_candidates(orb_extractor* ex, int level, int* xyr, int capacity, int* n);

#ifdef __cplusplus
}
#endif
#endif
