// orb_vocabulary.cpp -- C ABI of the DBoW2 vocabulary path (orb_gpu.h):
// the tree loaded from DBoW2's text format or from arrays, renumbered
// breadth-first into HBM, and the batched transform / BowVector /
// FeatureVector launches (kernels in orb_bow.hip).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "orb_internal.h"

using namespace orb_detail;

extern "C" hipError_t orb_launch_voc_transform(const VocDev* V, const uint8_t* desc, int n, int levelsup, int* word,
                                               double* weight, int* nid, hipStream_t st);
extern "C" hipError_t orb_launch_voc_bow(const VocDev* V, const BowJob* jobs, int n_images, const int* word,
                                         const double* weight, const int* nid, int* bow_word, double* bow_value,
                                         int* n_bow, int* fv_node, int* fv_feat, int* n_fv, hipStream_t st);

namespace {
const int kBowCap = 4096;          // descriptors per image in one BoW launch
const int kMaxChildren = 1 << 22;  // children of one node (kernel key field)

template <class T>
int grow(T** p, size_t* cap, size_t n)
{
    if (n <= *cap) return ORB_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, n * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return ORB_E_OOM;
    }
    *cap = n;
    return ORB_OK;
}
}  // namespace

struct orb_vocabulary {
    CtxRef ctx;
    int k = 0, L = 0, scoring = 0, weighting = 0, n_nodes = 0, n_words = 0, max_depth = 0;
    int2* d_meta = nullptr;
    uint4* d_desc = nullptr;
    int* d_orig = nullptr;
    int* d_word = nullptr;
    double* d_weight = nullptr;
    VocDev dev{};
    // per-call scratch, grown on demand
    uint8_t* d_in = nullptr;
    size_t in_cap = 0;
    int* d_rw = nullptr;
    size_t rw_cap = 0;
    double* d_rwt = nullptr;
    size_t rwt_cap = 0;
    int* d_rn = nullptr;
    size_t rn_cap = 0;
    int* d_bw = nullptr;
    size_t bw_cap = 0;
    double* d_bv = nullptr;
    size_t bv_cap = 0;
    int* d_fn = nullptr;
    size_t fn_cap = 0;
    int* d_ff = nullptr;
    size_t ff_cap = 0;
    BowJob* d_jobs = nullptr;
    size_t jobs_cap = 0;
    int* d_cnt = nullptr;
    size_t cnt_cap = 0;

    ~orb_vocabulary()
    {
        void* ps[] = {d_meta, d_desc, d_orig, d_word, d_weight, d_in, d_rw, d_rwt, d_rn,
                      d_bw, d_bv, d_fn, d_ff, d_jobs, d_cnt};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }

    // Tree from file-order arrays (node 0 = root, parent[i] < i), as
    // loadFromTextFile builds it (TemplatedVocabulary.h:1405-1465): children
    // in file order, word ids in file order of the word nodes; the root keeps
    // the Node() defaults (word 0, weight 0).
    int build(int n, const int32_t* parent, const uint8_t* is_word, const uint8_t* desc, const double* weight)
    {
        std::vector<std::vector<int>> ch(n);
        std::vector<int> word(n, 0);
        int nw = 0;
        for (int i = 1; i < n; i++) {
            if (parent[i] < 0 || parent[i] >= i) return ORB_E_ARG;
            ch[parent[i]].push_back(i);
            if (is_word[i] > 0) word[i] = nw++;
        }
        for (int i = 0; i < n; i++)
            if ((long long)ch[i].size() >= kMaxChildren) return ORB_E_ARG;
        // breadth-first renumbering: every node's children get consecutive ids
        std::vector<int> order;             // new id -> file id
        std::vector<int> depth(n, 0);
        order.reserve(n);
        order.push_back(0);
        std::vector<int2> meta(n);
        int md = 0;
        for (size_t q = 0; q < order.size(); q++) {
            const int f = order[q];
            meta[q] = make_int2((int)order.size(), (int)ch[f].size());
            for (int c : ch[f]) {
                depth[c] = depth[f] + 1;
                md = std::max(md, depth[c]);
                order.push_back(c);
            }
        }
        std::vector<uint8_t> hd((size_t)n * 32, 0);
        std::vector<int> horig(n), hword(n);
        std::vector<double> hw(n, 0.0);
        for (int q = 0; q < n; q++) {
            const int f = order[q];
            horig[q] = f;
            if (f == 0) {
                hword[q] = 0;
                hw[q] = 0.0;
            } else {
                std::memcpy(&hd[(size_t)q * 32], desc + (size_t)f * 32, 32);
                hword[q] = word[f];
                hw[q] = weight[f];
            }
        }
        if (hipMalloc(&d_meta, n * sizeof(int2)) != hipSuccess || hipMalloc(&d_desc, (size_t)n * 32) != hipSuccess ||
            hipMalloc(&d_orig, n * sizeof(int)) != hipSuccess || hipMalloc(&d_word, n * sizeof(int)) != hipSuccess ||
            hipMalloc(&d_weight, n * sizeof(double)) != hipSuccess)
            return ORB_E_OOM;
        HIP_TRY(hipMemcpy(d_meta, meta.data(), n * sizeof(int2), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_desc, hd.data(), (size_t)n * 32, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_orig, horig.data(), n * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_word, hword.data(), n * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d_weight, hw.data(), n * sizeof(double), hipMemcpyHostToDevice));
        n_nodes = n;
        n_words = nw;
        max_depth = md;
        dev.meta = d_meta;
        dev.desc = d_desc;
        dev.orig_id = d_orig;
        dev.word_id = d_word;
        dev.weight = d_weight;
        dev.n_nodes = n;
        dev.L = L;
        dev.max_depth = md;
        dev.scoring = scoring;
        dev.weighting = weighting;
        return ORB_OK;
    }

    int ensure_rows(size_t rows)
    {
        int rc = grow(&d_in, &in_cap, rows * 32);
        if (!rc) rc = grow(&d_rw, &rw_cap, rows);
        if (!rc) rc = grow(&d_rwt, &rwt_cap, rows);
        if (!rc) rc = grow(&d_rn, &rn_cap, rows);
        if (!rc) rc = grow(&d_bw, &bw_cap, rows);
        if (!rc) rc = grow(&d_bv, &bv_cap, rows);
        if (!rc) rc = grow(&d_fn, &fn_cap, rows);
        if (!rc) rc = grow(&d_ff, &ff_cap, rows);
        return rc;
    }

    // Stage a batch: descriptors and the per-image row ranges.
    int stage(int n_images, const uint8_t* desc, const int32_t* offsets, std::vector<BowJob>& jobs)
    {
        jobs.resize(n_images);
        for (int i = 0; i < n_images; i++) {
            const int a = offsets[i], b = offsets[i + 1];
            if (a < 0 || b < a) return ORB_E_ARG;
            if (b - a > kBowCap) return ORB_E_CAPACITY;
            jobs[i].off = a;
            jobs[i].n = b - a;
        }
        const size_t rows = (size_t)offsets[n_images];
        int rc = ensure_rows(std::max<size_t>(rows, 1));
        if (!rc) rc = grow(&d_jobs, &jobs_cap, (size_t)std::max(n_images, 1));
        if (!rc) rc = grow(&d_cnt, &cnt_cap, 2 * (size_t)std::max(n_images, 1));
        if (rc) return rc;
        hipStream_t st = ctx->stream;
        if (rows) HIP_TRY(hipMemcpyAsync(d_in, desc, rows * 32, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), n_images * sizeof(BowJob), hipMemcpyHostToDevice, st));
        return ORB_OK;
    }

    int enqueue_bow(int n_images, int rows, int levelsup)
    {
        hipStream_t st = ctx->stream;
        HIP_TRY(orb_launch_voc_transform(&dev, d_in, rows, levelsup, d_rw, d_rwt, d_rn, st));
        if (n_words == 0) {
            HIP_TRY(hipMemsetAsync(d_cnt, 0, 2 * n_images * sizeof(int), st));
            return ORB_OK;
        }
        HIP_TRY(orb_launch_voc_bow(&dev, d_jobs, n_images, d_rw, d_rwt, d_rn, d_bw, d_bv, d_cnt, d_fn, d_ff,
                                   d_cnt + n_images, st));
        return ORB_OK;
    }
};

namespace {

int finish_create(orb_vocabulary* v, int n, const int32_t* parent, const uint8_t* is_word, const uint8_t* desc,
                  const double* weight, orb_vocabulary** out)
{
    int rc = v->build(n, parent, is_word, desc, weight);
    if (rc) {
        delete v;
        return rc;
    }
    *out = v;
    return ORB_OK;
}

// Reads an int the way `stringstream >> int` does for well-formed text.
bool next_long(const char*& p, long* x)
{
    char* e = nullptr;
    errno = 0;
    const long v = std::strtol(p, &e, 10);
    if (e == p) return false;
    p = e;
    *x = v;
    return true;
}

}  // namespace

extern "C" {

int orb_vocabulary_create(orb_context* ctx, int k, int L, int scoring, int weighting, int n_nodes,
                          const int32_t* parent, const uint8_t* is_word, const uint8_t* desc, const double* weight,
                          orb_vocabulary** out)
{
    if (!ctx || !out || n_nodes < 1 || (n_nodes > 1 && (!parent || !is_word || !desc || !weight))) return ORB_E_ARG;
    if (scoring < 0 || scoring > 5 || weighting < 0 || weighting > 3) return ORB_E_ARG;
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    orb_vocabulary* v = new orb_vocabulary();
    v->ctx = ctx;
    v->k = k;
    v->L = L;
    v->scoring = scoring;
    v->weighting = weighting;
    return finish_create(v, n_nodes, parent, is_word, desc, weight, out);
}

// TemplatedVocabulary::loadFromTextFile (TemplatedVocabulary.h:1368-1470):
// header "k L scoring weighting" (k in [0,20], L in [1,10], scoring in [0,5],
// weighting in [0,3]); then per line: parent id, is-leaf flag, 32 descriptor
// values (ints, stored as unsigned char), weight.  Blank lines are skipped
// (the reference turns a trailing one into an extra root child with
// uninitialised descriptor bytes).
int orb_vocabulary_load_text(orb_context* ctx, const char* path, orb_vocabulary** out)
{
    if (!ctx || !path || !out) return ORB_E_ARG;
    *out = nullptr;
    FILE* f = std::fopen(path, "r");
    if (!f) return ORB_E_ARG;
    std::string line;
    auto getline = [&](std::string& s) {
        s.clear();
        int c;
        bool any = false;
        while ((c = std::fgetc(f)) != EOF) {
            any = true;
            if (c == '\n') break;
            s.push_back((char)c);
        }
        return any;
    };
    int rc = ORB_OK;
    std::vector<int32_t> parent{0};
    std::vector<uint8_t> is_word{0};
    std::vector<uint8_t> desc(32, 0);
    std::vector<double> weight{0.0};
    int k = 0, L = 0, sc = 0, wt = 0;
    if (!getline(line)) {
        rc = ORB_E_ARG;
    } else {
        long h[4] = {-1, -1, -1, -1};
        const char* p = line.c_str();
        for (int i = 0; i < 4; i++)
            if (!next_long(p, &h[i])) break;
        k = (int)h[0];
        L = (int)h[1];
        sc = (int)h[2];
        wt = (int)h[3];
        if (k < 0 || k > 20 || L < 1 || L > 10 || sc < 0 || sc > 5 || wt < 0 || wt > 3) rc = ORB_E_ARG;
    }
    while (!rc && getline(line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        const char* p = line.c_str();
        long pid = 0, leaf = 0;
        if (!next_long(p, &pid) || !next_long(p, &leaf)) {
            rc = ORB_E_ARG;
            break;
        }
        uint8_t d[32] = {0};
        for (int i = 0; i < 32; i++) {
            long x = 0;
            if (next_long(p, &x)) d[i] = (uint8_t)x;
        }
        char* e = nullptr;
        const double w = std::strtod(p, &e);
        const int nid = (int)parent.size();
        if (pid < 0 || pid >= nid) {
            rc = ORB_E_ARG;
            break;
        }
        parent.push_back((int32_t)pid);
        is_word.push_back(leaf > 0 ? 1 : 0);
        desc.insert(desc.end(), d, d + 32);
        weight.push_back(e == p ? 0.0 : w);
    }
    std::fclose(f);
    if (rc) return rc;
    return orb_vocabulary_create(ctx, k, L, sc, wt, (int)parent.size(), parent.data(), is_word.data(), desc.data(),
                                 weight.data(), out);
}

void orb_vocabulary_destroy(orb_vocabulary* v)
{
    if (!v) return;
    (void)hipSetDevice(v->ctx->device);
    (void)hipStreamSynchronize(v->ctx->stream);
    delete v;
}

int orb_vocabulary_info(const orb_vocabulary* v, int* k, int* L, int* scoring, int* weighting, int* n_nodes,
                        int* n_words)
{
    if (!v) return ORB_E_ARG;
    if (k) *k = v->k;
    if (L) *L = v->L;
    if (scoring) *scoring = v->scoring;
    if (weighting) *weighting = v->weighting;
    if (n_nodes) *n_nodes = v->n_nodes;
    if (n_words) *n_words = v->n_words;
    return ORB_OK;
}

int orb_vocabulary_transform(orb_vocabulary* v, const uint8_t* desc, int n, int levelsup, int32_t* word_id,
                             double* weight, int32_t* node_id)
{
    if (!v || n < 0 || (n && (!desc || !word_id || !weight))) return ORB_E_ARG;
    if (n == 0) return ORB_OK;
    HIP_TRY(hipSetDevice(v->ctx->device));
    int rc = v->ensure_rows((size_t)n);
    if (rc) return rc;
    hipStream_t st = v->ctx->stream;
    HIP_TRY(hipMemcpyAsync(v->d_in, desc, (size_t)n * 32, hipMemcpyHostToDevice, st));
    HIP_TRY(orb_launch_voc_transform(&v->dev, v->d_in, n, levelsup, v->d_rw, v->d_rwt, v->d_rn, st));
    HIP_TRY(hipMemcpyAsync(word_id, v->d_rw, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(weight, v->d_rwt, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    if (node_id) HIP_TRY(hipMemcpyAsync(node_id, v->d_rn, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

int orb_vocabulary_bow(orb_vocabulary* v, int n_images, const uint8_t* desc, const int32_t* offsets, int levelsup,
                       int32_t* bow_word, double* bow_value, int32_t* n_bow, int32_t* fv_node, int32_t* fv_feature,
                       int32_t* n_fv)
{
    if (!v || n_images < 0 || (n_images && (!offsets || !bow_word || !bow_value || !n_bow || !fv_node ||
                                            !fv_feature || !n_fv)))
        return ORB_E_ARG;
    if (n_images == 0) return ORB_OK;
    if (offsets[n_images] > 0 && !desc) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(v->ctx->device));
    std::vector<BowJob> jobs;
    int rc = v->stage(n_images, desc, offsets, jobs);
    if (rc) return rc;
    const int rows = offsets[n_images];
    rc = v->enqueue_bow(n_images, rows, levelsup);
    if (rc) return rc;
    hipStream_t st = v->ctx->stream;
    if (rows) {
        HIP_TRY(hipMemcpyAsync(bow_word, v->d_bw, (size_t)rows * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(bow_value, v->d_bv, (size_t)rows * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(fv_node, v->d_fn, (size_t)rows * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(fv_feature, v->d_ff, (size_t)rows * 4, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipMemcpyAsync(n_bow, v->d_cnt, n_images * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(n_fv, v->d_cnt + n_images, n_images * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return ORB_OK;
}

int orb_vocabulary_bow_time(orb_vocabulary* v, int n_images, const uint8_t* desc, const int32_t* offsets,
                            int levelsup, int steps, float* ms)
{
    if (!v || n_images <= 0 || !desc || !offsets || steps <= 0 || !ms) return ORB_E_ARG;
    HIP_TRY(hipSetDevice(v->ctx->device));
    std::vector<BowJob> jobs;
    int rc = v->stage(n_images, desc, offsets, jobs);
    if (rc) return rc;
    const int rows = offsets[n_images];
    hipStream_t st = v->ctx->stream;
    rc = v->enqueue_bow(n_images, rows, levelsup);     // warm-up
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    (void)hipEventRecord(e0, st);
    for (int s = 0; s < steps && !rc; s++) rc = v->enqueue_bow(n_images, rows, levelsup);
    (void)hipEventRecord(e1, st);
    const hipError_t se = hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (se != hipSuccess) return ORB_E_DEVICE;
    *ms = t / steps;
    return ORB_OK;
}

}  // extern "C"
