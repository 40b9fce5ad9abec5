// orb_types.h -- job descriptors shared by the host runtime and the gfx950
// kernels of the ORB front end.
//
// One ImageJob describes one image going through the extraction pipeline
// (pyramid -> FAST cells -> quadtree distribution -> orientation + rotated
// BRIEF).  Every job carries device pointers to its own buffers, so a single
// launch of each stage can cover any set of images: the two views of a stereo
// frame, a mixed-size batch, or the frames admitted by the serving loop.
#pragma once
#include <stdint.h>

#define ORB_MAX_LEVELS 16
#define ORB_DESC_KPB 16             // keypoints per describe workgroup (16 lanes each)
#define ORB_MAX_ROOTS 16
#define ORB_LIST_CAP 1024      // quadtree list capacity per level (LDS; one node per octree thread)
#define ORB_TILE_W 64          // level kernel output tile
#define ORB_EDGE 19            // EDGE_THRESHOLD, ORBextractor.cc:81
#define ORB_MIN_BORDER 16      // EDGE_THRESHOLD - 3, ORBextractor.cc:1049

// Counter slots per job (int32, zeroed at the start of each run).
#define ORB_CNT_KEYS 0                       // [16] FAST candidates per level
#define ORB_CNT_RET (ORB_CNT_KEYS + 16)      // [16] keypoints kept per level
#define ORB_CNT_NOUT (ORB_CNT_RET + 16)      // total keypoints
#define ORB_CNT_ERR (ORB_CNT_NOUT + 1)       // error bits
#define ORB_CNT_MATCH (ORB_CNT_ERR + 1)      // stereo matches recorded
#define ORB_CNT_SLOTS 64

#define ORB_KEY_RETRY (1ull << 40)   // candidate from the minThFAST pass

#define ORB_ERR_KEYS 1        // candidate buffer overflow
#define ORB_ERR_LIST 2        // quadtree list overflow
#define ORB_ERR_OUT 4         // output capacity overflow

// cv::KeyPoint layout (28 bytes).
struct OrbKeyPoint {
    float x, y, size, angle, response;
    int octave, class_id;
};

struct LevelJob {
    int w, h, pitch;                 // level size and row pitch (bytes)
    int tiles_x, tiles_y;            // level kernel tiles
    int nCols, nRows, wCell, hCell;  // FAST cell grid (ORBextractor.cc:1065-1071)
    unsigned wrcp, hrcp;             // ceil(2^32 / wCell), ceil(2^32 / hCell): x / cell = umulhi(x, rcp)
    int maxBorderX, maxBorderY;
    int quota;                       // mnFeaturesPerLevel[level]
    int nIni;                        // quadtree roots
    float hX;                        // root width (float, ORBextractor.cc:701)
    float scale;                     // mvScaleFactor[level]
    int root_x[ORB_MAX_ROOTS + 1];   // (int)(hX*i), i = 0..nIni
    int key_cap;                     // candidate capacity (both thresholds)
    int ncells;                      // nCols * nRows
    int list_cap;                    // kept-keypoint capacity (<= ORB_LIST_CAP)
    int rs_end;                      // level columns < rs_end: SSE2 resize rounding (else scalar)
    int bs_end;                      // level columns < bs_end: SSE2 blur-column rounding (else scalar)
    uint8_t* lvl;                    // level image
    uint8_t* blur;                   // 7x7 sigma-2 blur of the level
    uint8_t* cellflag;               // per FAST cell: has an iniThFAST survivor
    unsigned long long* keys;        // candidates: response<<32 | order key
    unsigned long long* keys2;       // candidates after the per-cell retry
    int* node;                       // per-candidate quadtree node (scratch)
    unsigned long long* ret;         // kept keypoints in output order
    const int2* xtab;                // levels >= 1: per column {sx | sse2 << 30, a0 | a1 << 16} of the
    const int2* ytab;                // resize from level-1; per row {sy, b0 | b1 << 16}
    const int2* xwin;                // levels >= 1: per tile column / row, the source
    const int2* ywin;                // columns / rows {first, last} the tile's taps read
};

struct ImageJob {
    const uint8_t* src;              // input image (level 0), src_pitch bytes per row; the
    int src_pitch;                   // level-0 kernel copies it into lv[0].lvl unless it is lv[0].lvl
    int nlevels, iniTh, minTh;
    int out_cap;                     // keypoint output capacity
    int* counters;                   // ORB_CNT_SLOTS ints
    OrbKeyPoint* kps;                // out_cap keypoints
    uint8_t* desc;                   // out_cap x 32 bytes
    LevelJob lv[ORB_MAX_LEVELS];
};

#define ORB_MAX_JOBS 256

// Launch table for one run over a set of jobs (prefix sums in block units).
struct RunTable {
    int njobs;
    int max_levels;
    int tile_start[ORB_MAX_LEVELS][ORB_MAX_JOBS + 1];   // per level: first tile block of job j
    int desc_start[(ORB_MAX_JOBS + 1) * ORB_MAX_LEVELS];  // describe kernel: first block of (job, level)
    int tiles_per_job[ORB_MAX_LEVELS];  // > 0: every job has this many tiles at the level
    int tile_h[ORB_MAX_LEVELS];         // level kernel tile height at the level (32 or 16)
    int desc_per_job;                   // > 0: every job has this many describe blocks
};

// Arguments of one stereo match (Frame::ComputeStereoMatches).
struct StereoArgs {
    const ImageJob* left;
    const ImageJob* right;
    const OrbKeyPoint* kl;
    const uint8_t* dl;
    const int* nl;
    const OrbKeyPoint* kr;
    const uint8_t* dr;
    const int* nr;
    int cap_l;
    int nrows;
    float bf, mb;
    float scale[ORB_MAX_LEVELS], inv_scale[ORB_MAX_LEVELS];
    float* uRight;
    float* depth;
    int* match_sad;         // per left keypoint: SAD of its match, or -1
    int4* srec;             // scratch: right keypoints grouped by row bucket as
                            // {minr, maxr, index << 4 | octave, x bits} ...
    uint8_t* sdesc;         // ... and their descriptors in the same order
    int2* rows;             // per (octave o, level-0 row): [first, end) of the entries
                            // of octave o whose row band can hold the row (nlev x nrows)
    int cap_r;
    int nlev;               // pyramid levels of both extractors
};

// Arguments of one brute-force match of a query set against a train set
// (serving: this frame's descriptors against the stream's previous ones).
struct MatchArgs {
    const uint8_t* q;
    const int* nq;
    const uint8_t* t;
    const int* nt;
    int th_low;
    float ratio;
    int* idx;
    int* dist;
    int cap_q;
    int wave_base;          // first wavefront of this pair in a multi-pair launch
    uint8_t* save_desc;     // after matching: q copied here ...
    int* save_n;            // ... and *nq here (the stream's resident state)
    int save_cap;
};

// One region of a batch's staged host transfers, copied in 4-byte words:
// `words` of them, or min(words, *count * unit_words) when count is set (a
// slot's keypoint count, read on the device).  Source and destination are
// 4-byte aligned; a region that is a whole number of 16-byte words at
// 16-byte aligned addresses is copied 16 bytes at a time.  Either side may
// be pinned host memory, which the device reads and writes over PCIe.
struct CopyOp {
    const unsigned* src;
    unsigned* dst;
    const int* count;
    int words;
    int unit_words;
};

// A DBoW2 vocabulary tree on the device, nodes renumbered breadth-first so
// that every node's children are contiguous (in their file order).
struct VocDev {
    const int2* meta;       // per node: first child, number of children
    const uint4* desc;      // per node: 32-byte descriptor as two uint4
    const int* orig_id;     // per node: id in the file (node 0 = root)
    const int* word_id;     // per node: word id (0 for non-words)
    const double* weight;   // per node: weight from the file
    int n_nodes;
    int L;                  // depth levels from the header (nid level = L - levelsup)
    int max_depth;          // deepest node, bounds every descent
    int scoring, weighting;
};

// One image of a BoW batch: its rows [off, off + n) of the descriptor and
// per-descriptor arrays; outputs at the same offsets.
struct BowJob {
    int off, n;
};

// One frame pair of ORBmatcher::SearchForInitialization: F1 / F2 keypoints
// (mvKeysUn) and descriptors, the grid bounds (mnMinX .. mnMaxY), the search
// positions (vbPrevMatched, updated in place), parameters, scratch and
// outputs.
struct WinPair {
    const OrbKeyPoint* kp1;
    const uint8_t* d1;
    const OrbKeyPoint* kp2;
    const uint8_t* d2;
    int n1, n2;
    float minX, maxX, minY, maxY;
    float2* prev;
    int window;
    float nnratio;
    int check_orientation;
    int* sorted;            // n2: F2 keypoints sorted by (cell, index)
    int* cell_start;        // 64*48 + 1
    int* cnt;               // n1: candidates per query
    int* off;               // n1 + 1: their offsets
    int2* cand;             // (i2, distance) lists
    int* m12;               // n1: vnMatches12
    int* nmatch;            // nmatches
    // SearchForInitialization's in-order pass as parallel rounds (unused by
    // the grid kernels of SearchByProjection, which leave them null)
    int* out[2];            // n1 per round parity: accepted target << 9 | distance, or -1
    unsigned* claim[3];     // n2 * ORB_WIN_SLOTS per round mod 3: query << 9 | distance
    int* ccnt[3];           // n2 per round mod 3: claims written (beyond ORB_WIN_SLOTS: overflow)
    int* changed;           // [ORB_WIN_MAXIT]: bit 0 an outcome changed, bit 1 claims overflowed,
                            // bit 2 the round has not run
};
#define ORB_WIN_MAXIT 48
#define ORB_WIN_SLOTS 8

// ORBmatcher::SearchByProjection: one query (a map point to place in the
// frame) after the per-point arithmetic of either overload.
struct ProjQuery {
    float x, y, r;          // window centre and radius (GetFeaturesInArea)
    float su;               // right-image u compared with mvuRight (radius r)
    int minL, maxL;         // GetFeaturesInArea level arguments
    int flags;              // ORB_PQ_VALID | ORB_PQ_OBS
    float angle;            // last-frame keypoint angle (rotation check)
};
#define ORB_PQ_VALID 1
#define ORB_PQ_OBS 4
#define ORB_PROJ_MAXIT 48   // claim-resolution rounds before the sequential pass

// One frame of a SearchByProjection call: its keypoints (mvKeysUn),
// descriptors, mvuRight, occupancy and grid, the queries, their candidate
// lists, and the claim-resolution state.
struct ProjFrame {
    const OrbKeyPoint* kp;
    const uint8_t* desc;
    const float* uright;
    const uint8_t* occupied;    // mvpMapPoints[i] && Observations() > 0
    int n;
    float minX, maxX, minY, maxY;
    const int* sorted;          // grid, as WinPair's
    const int* cell_start;
    ProjQuery* q;
    const uint8_t* qdesc;
    int nq;
    int* cnt;                   // candidates per query
    const int* off;             // their offsets (nq + 1)
    unsigned* cand;             // i2 | distance << 12 | octave << 21, GetFeaturesInArea order
    int* out[2];                // outcome per query by round parity: keypoint or -1
    int* claim[3];              // per keypoint: first claiming query (with observations), by round mod 3
    int* changed;               // [ORB_PROJ_MAXIT]: bit 0 round k changed some outcome, bit 2 not run
    int rule;                   // 1: local map (best + second, level ratio); 2: last frame (best, rotation)
    float nnratio;
    int check_orientation;
    int* pt_match;              // outputs
    int* key_point;
    int* nmatch;
};
