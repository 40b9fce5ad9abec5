// bow_oracle.cpp -- CPU oracle for the DBoW2 vocabulary transform.
//
// TEST INFRASTRUCTURE ONLY (see orb_oracle.h): the checker for the GPU
// vocabulary path and the CPU baseline that bench.py times.
//
// Restates, from Thirdparty/DBoW2/DBoW2:
//   TemplatedVocabulary::loadFromTextFile   TemplatedVocabulary.h:1368-1470
//   TemplatedVocabulary::saveToTextFile     TemplatedVocabulary.h:1475-1497
//   TemplatedVocabulary::transform(feature, word_id, weight, nid, levelsup)
//                                           TemplatedVocabulary.h:1240-1290
//   TemplatedVocabulary::transform(features, BowVector, FeatureVector,
//                                  levelsup) TemplatedVocabulary.h:1136-1206
//   BowVector::addWeight / addIfNotExist / normalize   BowVector.cpp
//   FeatureVector::addFeature                          FeatureVector.cpp:37-53
//   FORB::distance (popcount of XOR)                   FORB.cpp:81-101
//   ScoringObject mustNormalize per scoring type       ScoringObject.h
// and MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cc:365-450),
// with the same containers (std::map) so that iteration and summation order
// are the reference's.
//
// One deliberate difference: loadFromTextFile's `while(!f.eof())` turns a
// trailing empty line into an extra child of the root whose descriptor bytes
// are never written (uninitialised memory).  Blank lines are skipped here.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "orb_oracle.h"

namespace {

struct VNode {
    int parent = 0;
    std::vector<int> children;
    uint8_t desc[32] = {0};
    double weight = 0.0;
    int word_id = 0;          // Node() default for non-word nodes
};

struct Voc {
    int k = 0, L = 0, scoring = 0, weighting = 0;
    std::vector<VNode> nodes;
    std::vector<int> words;   // word id -> node id
};

// FORB::distance: popcount of the XOR of 8 32-bit words.
int forb_distance(const uint8_t* a, const uint8_t* b)
{
    int d = 0;
    for (int i = 0; i < 32; i++) d += __builtin_popcount((unsigned)(a[i] ^ b[i]));
    return d;
}

// ScoringObject.h: L1_NORM, L2_NORM, CHI_SQUARE, KL, BHATTACHARYYA normalise
// (L2 for L2_NORM, L1 otherwise); DOT_PRODUCT does not.
bool must_normalize(int scoring, bool* l2)
{
    *l2 = scoring == 1;
    return scoring != 5;
}

void add_node(Voc& v, int pid, int is_word, const uint8_t* desc, double weight)
{
    const int nid = (int)v.nodes.size();
    v.nodes.emplace_back();
    VNode& n = v.nodes.back();
    n.parent = pid;
    std::memcpy(n.desc, desc, 32);
    n.weight = weight;
    v.nodes[pid].children.push_back(nid);
    if (is_word > 0) {
        v.nodes[nid].word_id = (int)v.words.size();
        v.words.push_back(nid);
    }
}

}  // namespace

extern "C" {

oracle_vocabulary* oracle_voc_create(int k, int L, int scoring, int weighting, int n_nodes, const int32_t* parent,
                                     const uint8_t* is_word, const uint8_t* desc, const double* weight)
{
    if (n_nodes < 1) return nullptr;
    Voc* v = new Voc();
    v->k = k;
    v->L = L;
    v->scoring = scoring;
    v->weighting = weighting;
    v->nodes.resize(1);
    for (int i = 1; i < n_nodes; i++) {
        if (parent[i] < 0 || parent[i] >= i) {
            delete v;
            return nullptr;
        }
        add_node(*v, parent[i], is_word[i], desc + (size_t)i * 32, weight[i]);
    }
    return reinterpret_cast<oracle_vocabulary*>(v);
}

// loadFromTextFile: header "k L scoring weighting" with the reference's range
// checks, then one line per node: parent, is-leaf flag, 32 descriptor bytes
// (parsed as int, stored as unsigned char), weight (double).
oracle_vocabulary* oracle_voc_load_text(const char* path)
{
    std::ifstream f(path);
    if (!f.is_open()) return nullptr;
    std::string s;
    if (!std::getline(f, s)) return nullptr;
    std::stringstream ss(s);
    int k = -1, L = -1, n1 = -1, n2 = -1;
    ss >> k >> L >> n1 >> n2;
    if (k < 0 || k > 20 || L < 1 || L > 10 || n1 < 0 || n1 > 5 || n2 < 0 || n2 > 3) return nullptr;
    Voc* v = new Voc();
    v->k = k;
    v->L = L;
    v->scoring = n1;
    v->weighting = n2;
    v->nodes.resize(1);
    std::string line;
    while (std::getline(f, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::stringstream sn(line);
        int pid = 0, leaf = 0;
        sn >> pid >> leaf;
        uint8_t d[32] = {0};
        for (int i = 0; i < 32; i++) {
            int x;
            sn >> x;
            if (!sn.fail()) d[i] = (uint8_t)x;
        }
        double w = 0.0;
        sn >> w;
        if (pid < 0 || pid >= (int)v->nodes.size()) {
            delete v;
            return nullptr;
        }
        add_node(*v, pid, leaf, d, w);
    }
    return reinterpret_cast<oracle_vocabulary*>(v);
}

// saveToTextFile: "k L  scoring weighting", then per node: parent, isLeaf()
// (= no children), each descriptor byte as an int followed by a space, a
// space, and the weight through a default-formatted ostream.
int oracle_voc_save_text(const oracle_vocabulary* voc, const char* path)
{
    const Voc& v = *reinterpret_cast<const Voc*>(voc);
    std::ofstream f(path);
    if (!f.is_open()) return -1;
    f << v.k << " " << v.L << " " << " " << v.scoring << " " << v.weighting << std::endl;
    for (size_t i = 1; i < v.nodes.size(); i++) {
        const VNode& n = v.nodes[i];
        f << n.parent << " " << (n.children.empty() ? 1 : 0) << " ";
        std::stringstream sd;
        for (int b = 0; b < 32; b++) sd << (int)n.desc[b] << " ";
        f << sd.str() << " " << (double)n.weight << std::endl;
    }
    return 0;
}

void oracle_voc_destroy(oracle_vocabulary* voc) { delete reinterpret_cast<Voc*>(voc); }

int oracle_voc_info(const oracle_vocabulary* voc, int* k, int* L, int* scoring, int* weighting, int* n_nodes,
                    int* n_words)
{
    const Voc& v = *reinterpret_cast<const Voc*>(voc);
    if (k) *k = v.k;
    if (L) *L = v.L;
    if (scoring) *scoring = v.scoring;
    if (weighting) *weighting = v.weighting;
    if (n_nodes) *n_nodes = (int)v.nodes.size();
    if (n_words) *n_words = (int)v.words.size();
    return 0;
}

// Node table in id order: parent, word id (0 for non-words), descriptor,
// weight -- what a loader produced, for comparison.
int oracle_voc_nodes(const oracle_vocabulary* voc, int32_t* parent, int32_t* word_id, uint8_t* desc, double* weight)
{
    const Voc& v = *reinterpret_cast<const Voc*>(voc);
    for (size_t i = 0; i < v.nodes.size(); i++) {
        if (parent) parent[i] = v.nodes[i].parent;
        if (word_id) word_id[i] = v.nodes[i].word_id;
        if (desc) std::memcpy(desc + i * 32, v.nodes[i].desc, 32);
        if (weight) weight[i] = v.nodes[i].weight;
    }
    return 0;
}

// transform(feature, word_id, weight, nid, levelsup), TemplatedVocabulary.h
// :1240-1290: from the root, take the child at the smallest distance (the
// first one on ties), until a node without children; nid is the node chosen
// at depth L - levelsup (the root when that is <= 0).
void oracle_voc_transform_one(const oracle_vocabulary* voc, const uint8_t* f, int levelsup, int* word_id,
                              double* weight, int* nid)
{
    const Voc& v = *reinterpret_cast<const Voc*>(voc);
    const int nid_level = v.L - levelsup;
    if (nid_level <= 0 && nid) *nid = 0;
    int final_id = 0, current_level = 0;
    if (v.nodes[0].children.empty()) {       // a vocabulary of the root alone
        *word_id = v.nodes[0].word_id;
        *weight = v.nodes[0].weight;
        return;
    }
    do {
        ++current_level;
        const std::vector<int>& ch = v.nodes[final_id].children;
        final_id = ch[0];
        double best_d = forb_distance(f, v.nodes[final_id].desc);
        for (size_t j = 1; j < ch.size(); j++) {
            const double d = forb_distance(f, v.nodes[ch[j]].desc);
            if (d < best_d) {
                best_d = d;
                final_id = ch[j];
            }
        }
        if (nid && current_level == nid_level) *nid = final_id;
    } while (!v.nodes[final_id].children.empty());
    *word_id = v.nodes[final_id].word_id;
    *weight = v.nodes[final_id].weight;
}

int oracle_voc_transform(const oracle_vocabulary* voc, const uint8_t* desc, int n, int levelsup, int32_t* word_id,
                         double* weight, int32_t* node_id)
{
    for (int i = 0; i < n; i++) {
        int w = 0, nid = 0;
        double wt = 0.0;
        oracle_voc_transform_one(voc, desc + (size_t)i * 32, levelsup, &w, &wt, &nid);
        word_id[i] = w;
        weight[i] = wt;
        if (node_id) node_id[i] = nid;
    }
    return 0;
}

// transform(features, BowVector, FeatureVector, levelsup),
// TemplatedVocabulary.h:1136-1206.  Outputs in map order: BowVector as
// (word id, value) sorted by word id; FeatureVector as (node id, feature
// index) sorted by node id, features in insertion (= index) order.
int oracle_voc_bow(const oracle_vocabulary* voc, const uint8_t* desc, int n, int levelsup, int32_t* bow_word,
                   double* bow_value, int* n_bow, int32_t* fv_node, int32_t* fv_feature, int* n_fv)
{
    const Voc& v = *reinterpret_cast<const Voc*>(voc);
    std::map<unsigned, double> bow;
    std::map<unsigned, std::vector<unsigned>> fv;
    *n_bow = 0;
    *n_fv = 0;
    if (v.words.empty()) return 0;
    bool l2 = false;
    const bool must = must_normalize(v.scoring, &l2);
    const bool tf = v.weighting == 0 || v.weighting == 1;     // TF_IDF, TF
    for (int i = 0; i < n; i++) {
        int id = 0, nid = 0;
        double w = 0.0;
        oracle_voc_transform_one(voc, desc + (size_t)i * 32, levelsup, &id, &w, &nid);
        if (w > 0) {
            auto it = bow.lower_bound((unsigned)id);
            if (it != bow.end() && it->first == (unsigned)id) {
                if (tf) it->second += w;                         // addWeight
            } else {
                bow.insert(it, {(unsigned)id, w});               // addWeight / addIfNotExist
            }
            fv[(unsigned)nid].push_back((unsigned)i);            // addFeature
        }
    }
    if (tf && !bow.empty() && !must) {
        const double nd = (double)bow.size();
        for (auto& e : bow) e.second /= nd;
    }
    if (must) {                                                  // BowVector::normalize
        double norm = 0.0;
        if (!l2) {
            for (auto& e : bow) norm += std::fabs(e.second);
        } else {
            for (auto& e : bow) norm += e.second * e.second;
            norm = std::sqrt(norm);
        }
        if (norm > 0.0)
            for (auto& e : bow) e.second /= norm;
    }
    int nb = 0;
    for (auto& e : bow) {
        bow_word[nb] = (int32_t)e.first;
        bow_value[nb] = e.second;
        nb++;
    }
    int nf = 0;
    for (auto& e : fv)
        for (unsigned f : e.second) {
            fv_node[nf] = (int32_t)e.first;
            fv_feature[nf] = (int32_t)f;
            nf++;
        }
    *n_bow = nb;
    *n_fv = nf;
    return 0;
}

// MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cc:365-450) for
// many points: point p owns rows offsets[p] .. offsets[p+1]-1 (its
// observations in the caller's order).  The full distance matrix, each row
// sorted, median = row[0.5*(N-1)] (index truncated), the first row with the
// smallest median wins.  best[p] is that row's index within the point's
// rows, or -1 for a point without rows.
int oracle_distinctive_descriptors(const uint8_t* desc, const int32_t* offsets, int n_points, int32_t* best)
{
    for (int p = 0; p < n_points; p++) {
        const int a = offsets[p], N = offsets[p + 1] - offsets[p];
        if (N <= 0) {
            best[p] = -1;
            continue;
        }
        std::vector<std::vector<float>> D(N, std::vector<float>(N, 0.f));
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++) {
                const int d = forb_distance(desc + (size_t)(a + i) * 32, desc + (size_t)(a + j) * 32);
                D[i][j] = (float)d;
                D[j][i] = (float)d;
            }
        int bestMedian = 0x7FFFFFFF, bestIdx = 0;
        for (int i = 0; i < N; i++) {
            std::vector<int> v(D[i].begin(), D[i].end());
            std::sort(v.begin(), v.end());
            const int median = v[(size_t)(0.5 * (N - 1))];
            if (median < bestMedian) {
                bestMedian = median;
                bestIdx = i;
            }
        }
        best[p] = bestIdx;
    }
    return 0;
}

// cvtColor(CV_{RGB,BGR,RGBA,BGRA}2GRAY) for 8-bit images, OpenCV 2.4/3.x
// (absent here) as ORB-SLAM2 calls it in Tracking::GrabImage*
// (src/Tracking.cc:250-276): 14-bit fixed-point weights R 4899, G 9617,
// B 1868 (round(16384 * {0.299, 0.587, 0.114})), rounding by + 2^13.
// bidx: the channel holding blue (0 for BGR/BGRA, 2 for RGB/RGBA).
int oracle_cvt_gray(const uint8_t* src, int w, int h, int stride, int scn, int bidx, uint8_t* dst)
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t* p = src + (size_t)y * stride + (size_t)x * scn;
            const int b = p[bidx], g = p[1], r = p[bidx ^ 2];
            dst[(size_t)y * w + x] = (uint8_t)((b * 1868 + g * 9617 + r * 4899 + (1 << 13)) >> 14);
        }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- rectification
// cv::initUndistortRectifyMap(K, D, R, P, size, CV_32FC1, map1, map2) and
// cv::remap(src, dst, map1, map2, INTER_LINEAR) with the default constant
// zero border, as Examples/Stereo/stereo_euroc.cc:124-188 calls them
// (OpenCV 2.4/3.x, absent here, restated from its published algorithm).
namespace {

// Matx33d::inv(DECOMP_LU) for 3x3: adjugate over the determinant.
bool inv3(const double a[9], double b[9])
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

// D holds nD coefficients in OpenCV's order k1 k2 p1 p2 [k3 [k4 k5 k6 [s1 s2
// s3 s4]]] (nD = 0, 4, 5, 8 or 12; the tilted model is not restated).  R may
// be NULL (identity).  P is the 3x3 new camera matrix.
int oracle_init_rectify_map(const double* K, const double* D, int nD, const double* R, const double* P, int w, int h,
                            float* mapx, float* mapy)
{
    if (!(nD == 0 || nD == 4 || nD == 5 || nD == 8 || nD == 12)) return -1;
    double c[12] = {0};
    for (int i = 0; i < nD; i++) c[i] = D[i];
    const double k1 = c[0], k2 = c[1], p1 = c[2], p2 = c[3], k3 = c[4], k4 = c[5], k5 = c[6], k6 = c[7];
    const double s1 = c[8], s2 = c[9], s3 = c[10], s4 = c[11];
    double Rm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (R)
        for (int i = 0; i < 9; i++) Rm[i] = R[i];
    double PR[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) PR[i * 3 + j] = P[i * 3] * Rm[j] + P[i * 3 + 1] * Rm[3 + j] + P[i * 3 + 2] * Rm[6 + j];
    double ir[9];
    if (!inv3(PR, ir)) return -1;
    const double u0 = K[2], v0 = K[5], fx = K[0], fy = K[4];
    for (int i = 0; i < h; i++) {
        double _x = i * ir[1] + ir[2], _y = i * ir[4] + ir[5], _w = i * ir[7] + ir[8];
        for (int j = 0; j < w; j++, _x += ir[0], _y += ir[3], _w += ir[6]) {
            const double ww = 1. / _w, x = _x * ww, y = _y * ww;
            const double x2 = x * x, y2 = y * y;
            const double r2 = x2 + y2, _2xy = 2 * x * y;
            const double kr = (1 + ((k3 * r2 + k2) * r2 + k1) * r2) / (1 + ((k6 * r2 + k5) * r2 + k4) * r2);
            const double xd = (x * kr + p1 * _2xy + p2 * (r2 + 2 * x2) + s1 * r2 + s2 * r2 * r2);
            const double yd = (y * kr + p1 * (r2 + 2 * y2) + p2 * _2xy + s3 * r2 + s4 * r2 * r2);
            // the identity tilt: vecTilt = (xd, yd, 1), invProj = 1
            const double u = fx * 1.0 * xd + u0;
            const double v = fy * 1.0 * yd + v0;
            mapx[(size_t)i * w + j] = (float)u;
            mapy[(size_t)i * w + j] = (float)v;
        }
    }
    return 0;
}

// remap INTER_LINEAR, 8-bit, border constant 0: X = cvRound(mapx*32),
// Y = cvRound(mapy*32); taps at (X>>5, Y>>5), weights from the 32x32 table of
// products (1-fx)(1-fy), fx(1-fy), (1-fx)fy, fx fy scaled by 2^15 (exact
// integers here), out = (sum + 2^14) >> 15; taps outside the source read 0.
int oracle_remap_linear(const uint8_t* src, int sw, int sh, int sstride, const float* mapx, const float* mapy, int w,
                        int h, uint8_t* dst)
{
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
            const float fxm = mapx[(size_t)i * w + j] * 32.f, fym = mapy[(size_t)i * w + j] * 32.f;
            const double cx = std::min(std::max((double)fxm, -2147483648.0), 2147483647.0);
            const double cy = std::min(std::max((double)fym, -2147483648.0), 2147483647.0);
            const int X = (int)std::nearbyint(cx), Y = (int)std::nearbyint(cy);
            const int sx = std::min(std::max(X >> 5, -32768), 32767), sy = std::min(std::max(Y >> 5, -32768), 32767);
            const int ax = X & 31, ay = Y & 31;
            const int w00 = (32 - ay) * (32 - ax) * 32, w01 = (32 - ay) * ax * 32;
            const int w10 = ay * (32 - ax) * 32, w11 = ay * ax * 32;
            uint8_t out;
            if (sx >= sw || sx + 1 < 0 || sy >= sh || sy + 1 < 0) {
                out = 0;
            } else {
                auto at = [&](int x, int y) -> int {
                    return (x >= 0 && x < sw && y >= 0 && y < sh) ? src[(size_t)y * sstride + x] : 0;
                };
                const int v = at(sx, sy) * w00 + at(sx + 1, sy) * w01 + at(sx, sy + 1) * w10 + at(sx + 1, sy + 1) * w11;
                const int r = (v + (1 << 14)) >> 15;
                out = (uint8_t)std::min(std::max(r, 0), 255);
            }
            dst[(size_t)i * w + j] = out;
        }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- Frame keypoint ops
// Frame::UndistortKeyPoints (src/Frame.cc:724-776) and
// Frame::ComputeStereoFromRGBD (src/Frame.cc:1085-1123).  The first calls
// cv::undistortPoints(pts, pts, K, D, Mat(), K) when k1 != 0, restated from
// OpenCV 2.4/3.x's published cvUndistortPoints (absent here): in double,
// x = (u - cx) * (1/fx), five fixed-point iterations
//   icdist = 1 / (1 + ((k3 r2 + k2) r2 + k1) r2),
//   x = (x0 - (2 p1 x y + p2 (r2 + 2 x^2))) icdist,
//   y = (y0 - (p1 (r2 + 2 y^2) + 2 p2 x y)) icdist,
// then the new camera matrix P = K (with R = I) maps back: u = fx x + cx, cast
// to float.  (3.x's extra terms for k4-k6, s1-s4 and the tilt are exact zeros
// and ones for a 4- or 5-coefficient model, so both versions give these bits.)
extern "C" {

// xy: n (x, y) pairs; K4 = fx, fy, cx, cy; D: nD = 4 or 5 coefficients
// k1 k2 p1 p2 [k3].  k1 == 0 copies, as the reference does.
int oracle_undistort_points(const float* xy, int n, const float* K4, const float* D, int nD, float* out)
{
    if (!(nD == 4 || nD == 5) || n < 0) return -1;
    if (D[0] == 0.0f) {
        for (int i = 0; i < 2 * n; i++) out[i] = xy[i];
        return 0;
    }
    double k[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < nD; i++) k[i] = D[i];
    const double fx = K4[0], fy = K4[1], cx = K4[2], cy = K4[3];
    const double ifx = 1. / fx, ify = 1. / fy;
    for (int i = 0; i < n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        x = (x - cx) * ifx;
        y = (y - cy) * ify;
        const double x0 = x, y0 = y;
        for (int j = 0; j < 5; j++) {
            const double r2 = x * x + y * y;
            const double icdist = 1. / (1 + ((k[4] * r2 + k[1]) * r2 + k[0]) * r2);
            const double deltaX = 2 * k[2] * x * y + k[3] * (r2 + 2 * x * x);
            const double deltaY = k[2] * (r2 + 2 * y * y) + 2 * k[3] * x * y;
            x = (x0 - deltaX) * icdist;
            y = (y0 - deltaY) * icdist;
        }
        out[2 * i] = (float)(fx * x + cx);
        out[2 * i + 1] = (float)(fy * y + cy);
    }
    return 0;
}

// xy: the keypoints as extracted, xun their undistorted x; depth: h rows of
// w floats, `stride` floats apart.  imDepth.at<float>(v, u) takes the float
// coordinates as ints (truncation).  A keypoint outside the image (which the
// reference would read out of bounds) gets -1.
int oracle_rgbd_depth(const float* xy, const float* xun, int n, const float* depth, int w, int h, int stride,
                      float bf, float* uR, float* dep)
{
    for (int i = 0; i < n; i++) {
        uR[i] = -1;
        dep[i] = -1;
        const float v = xy[2 * i + 1], u = xy[2 * i];
        if (!(v >= 0 && u >= 0)) continue;
        const int r = (int)v, c = (int)u;
        if (r >= h || c >= w) continue;
        const float d = depth[(size_t)r * stride + c];
        if (d > 0) {
            dep[i] = d;
            uR[i] = xun[i] - bf / d;
        }
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- windowed matching
// ORBmatcher::SearchForInitialization (src/ORBmatcher.cc:573-717) with the
// keypoint grid of Frame::AssignFeaturesToGrid / PosInGrid /
// GetFeaturesInArea (src/Frame.cc:399-423, 590-700; this fork tests a circle,
// distx*distx + disty*disty < r*r) and ORBmatcher::ComputeThreeMaxima
// (:2035-2077).  kp arrays are mvKeysUn; bounds = mnMinX, mnMaxX, mnMinY,
// mnMaxY; prev (2 floats per F1 keypoint) is vbPrevMatched, updated in place.
namespace {

const int kGridCols = 64, kGridRows = 48, kHisto = 30;

struct Grid {
    std::vector<std::vector<int>> cell;   // [ix * rows + iy] -> keypoint indices in index order
    float minX, minY, invW, invH;
};

Grid build_grid(const oracle_keypoint* kp, int n, const float* b)
{
    Grid g;
    g.cell.assign(kGridCols * kGridRows, {});
    g.minX = b[0];
    g.minY = b[2];
    g.invW = static_cast<float>(kGridCols) / static_cast<float>(b[1] - b[0]);
    g.invH = static_cast<float>(kGridRows) / static_cast<float>(b[3] - b[2]);
    for (int i = 0; i < n; i++) {
        const int px = (int)std::round((kp[i].x - g.minX) * g.invW);
        const int py = (int)std::round((kp[i].y - g.minY) * g.invH);
        if (px < 0 || px >= kGridCols || py < 0 || py >= kGridRows) continue;
        g.cell[px * kGridRows + py].push_back(i);
    }
    return g;
}

std::vector<int> features_in_area(const Grid& g, const oracle_keypoint* kp, float x, float y, float r, int minLevel,
                                  int maxLevel)
{
    std::vector<int> out;
    const int nMinCellX = std::max(0, (int)std::floor((x - g.minX - r) * g.invW));
    if (nMinCellX >= kGridCols) return out;
    const int nMaxCellX = std::min(kGridCols - 1, (int)std::ceil((x - g.minX + r) * g.invW));
    if (nMaxCellX < 0) return out;
    const int nMinCellY = std::max(0, (int)std::floor((y - g.minY - r) * g.invH));
    if (nMinCellY >= kGridRows) return out;
    const int nMaxCellY = std::min(kGridRows - 1, (int)std::ceil((y - g.minY + r) * g.invH));
    if (nMaxCellY < 0) return out;
    const bool bCheckLevels = (minLevel > 0) || (maxLevel >= 0);
    for (int ix = nMinCellX; ix <= nMaxCellX; ix++)
        for (int iy = nMinCellY; iy <= nMaxCellY; iy++)
            for (int i : g.cell[ix * kGridRows + iy]) {
                if (bCheckLevels) {
                    if (kp[i].octave < minLevel) continue;
                    if (maxLevel >= 0 && kp[i].octave > maxLevel) continue;
                }
                const float distx = kp[i].x - x;
                const float disty = kp[i].y - y;
                if (oracle_planted(4) == 2) {     // planted: the original square window
                    if (std::fabs(distx) < r && std::fabs(disty) < r) out.push_back(i);
                    continue;
                }
                if (distx * distx + disty * disty < r * r) out.push_back(i);
            }
    return out;
}

void three_maxima(const std::vector<int>* histo, int L, int& ind1, int& ind2, int& ind3)
{
    int max1 = 0, max2 = 0, max3 = 0;
    for (int i = 0; i < L; i++) {
        const int s = (int)histo[i].size();
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
}

}  // namespace

extern "C" {

int oracle_search_for_initialization(const oracle_keypoint* kp1, const uint8_t* d1, int n1,
                                     const oracle_keypoint* kp2, const uint8_t* d2, int n2, const float* bounds,
                                     float* prev, int windowSize, float nnratio, int checkOrientation,
                                     int32_t* matches12)
{
    const int TH_LOW = 50;
    int nmatches = 0;
    for (int i = 0; i < n1; i++) matches12[i] = -1;
    const Grid g = build_grid(kp2, n2, bounds);
    std::vector<int> rotHist[kHisto];
    const float factor = kHisto / 360.0f;
    std::vector<int> vMatchedDistance(n2, 0x7FFFFFFF), vnMatches21(n2, -1);
    for (int i1 = 0; i1 < n1; i1++) {
        const int level1 = kp1[i1].octave;
        if (level1 > 0) continue;
        const std::vector<int> v2 =
            features_in_area(g, kp2, prev[2 * i1], prev[2 * i1 + 1], (float)windowSize, level1, level1);
        if (v2.empty()) continue;
        int bestDist = 0x7FFFFFFF, bestDist2 = 0x7FFFFFFF, bestIdx2 = -1;
        for (int i2 : v2) {
            const int dist = forb_distance(d1 + (size_t)i1 * 32, d2 + (size_t)i2 * 32);
            if (vMatchedDistance[i2] <= dist) continue;
            if (dist < bestDist) {
                bestDist2 = bestDist;
                bestDist = dist;
                bestIdx2 = i2;
            } else if (dist < bestDist2) {
                bestDist2 = dist;
            }
        }
        if (bestDist <= TH_LOW && bestDist < (float)bestDist2 * nnratio) {
            if (vnMatches21[bestIdx2] >= 0) {
                matches12[vnMatches21[bestIdx2]] = -1;
                nmatches--;
            }
            matches12[i1] = bestIdx2;
            vnMatches21[bestIdx2] = i1;
            vMatchedDistance[bestIdx2] = bestDist;
            nmatches++;
            if (checkOrientation) {
                float rot = kp1[i1].angle - kp2[bestIdx2].angle;
                if (rot < 0.0) rot += 360.0f;
                int bin = (int)std::round(rot * factor);
                if (bin == kHisto) bin = 0;
                rotHist[bin].push_back(i1);
            }
        }
    }
    if (checkOrientation) {
        int ind1 = -1, ind2 = -1, ind3 = -1;
        three_maxima(rotHist, kHisto, ind1, ind2, ind3);
        for (int i = 0; i < kHisto; i++) {
            if (i == ind1 || i == ind2 || i == ind3) continue;
            for (int idx1 : rotHist[i])
                if (matches12[idx1] >= 0) {
                    matches12[idx1] = -1;
                    nmatches--;
                }
        }
    }
    for (int i1 = 0; i1 < n1; i1++)
        if (matches12[i1] >= 0) {
            prev[2 * i1] = kp2[matches12[i1]].x;
            prev[2 * i1 + 1] = kp2[matches12[i1]].y;
        }
    return nmatches;
}

}  // extern "C"

// ---------------------------------------------------------- SearchByProjection
// Both per-frame overloads of ORBmatcher::SearchByProjection, statement by
// statement, with the same grid and GetFeaturesInArea as above.  The
// MapPoint / Frame members they read arrive as plain arrays; the pose algebra
// (isInFrustum for the first, Rcw * x3Dw + tcw and tlc for the second) stays
// with the caller.  Planted mistakes (family 4) are marked where they act.
namespace {

const int TH_HIGH = 100;

// ORBmatcher::RadiusByViewingCos (:186-202): the float viewCos is compared
// with the double 0.998.
float radius_by_viewing_cos(float viewCos)
{
    if (oracle_planted(4) == 1) return viewCos > 0.998f ? 2.5f : 4.0f;   // planted: float compare
    return viewCos > 0.998 ? 2.5f : 4.0f;
}

}  // namespace

extern "C" {

int oracle_search_by_projection(const oracle_keypoint* kp, const uint8_t* desc, const float* uright,
                                const uint8_t* occupied, int n, const float* bounds, const float* scale,
                                const oracle_proj_point* pts, const uint8_t* pdesc, int npts, float th,
                                float nnratio, int32_t* pt_match, int32_t* key_point)
{
    const int plant = oracle_planted(4);
    int nmatches = 0;
    const bool bFactor = th != 1.0;
    const Grid g = build_grid(kp, n, bounds);
    // mvpMapPoints[i]: -1 none, -2 a pre-existing point with observations
    // (occupied), else the index of the point that claimed it here
    std::vector<int> mp(n, -1);
    for (int i = 0; i < n; i++)
        if (occupied[i]) mp[i] = -2;
    auto has_obs = [&](int i) { return mp[i] == -2 || (mp[i] >= 0 && (pts[mp[i]].flags & 4)); };
    for (int i = 0; i < n; i++) key_point[i] = -1;
    for (int p = 0; p < npts; p++) {
        pt_match[p] = -1;
        const oracle_proj_point& P = pts[p];
        if (!(P.flags & 1)) continue;
        if (P.flags & 2) continue;
        const int nPredictedLevel = P.level;
        float r = radius_by_viewing_cos(P.view_cos);
        if (bFactor) r *= th;
        const std::vector<int> vIndices =
            features_in_area(g, kp, P.proj_x, P.proj_y, r * scale[nPredictedLevel], nPredictedLevel - 1,
                             nPredictedLevel);
        if (vIndices.empty()) continue;
        const uint8_t* MPdescriptor = pdesc + (size_t)p * 32;
        int bestDist = 256, bestLevel = -1, bestDist2 = 256, bestLevel2 = -1, bestIdx = -1;
        for (int idx : vIndices) {
            if (mp[idx] != -1 && plant != 4)              // planted 4: claims here never block
                if (has_obs(idx)) continue;
            if (plant == 4 && mp[idx] == -2) continue;
            if (uright[idx] > 0) {
                const float er = std::fabs(P.proj_xr - uright[idx]);
                if (plant == 5 ? er >= r * scale[nPredictedLevel] : er > r * scale[nPredictedLevel]) continue;
            }
            const int dist = forb_distance(MPdescriptor, desc + (size_t)idx * 32);
            if (dist < bestDist) {
                bestDist2 = bestDist;
                bestDist = dist;
                bestLevel2 = bestLevel;
                bestLevel = kp[idx].octave;
                bestIdx = idx;
            } else if (dist < bestDist2) {
                bestLevel2 = kp[idx].octave;
                bestDist2 = dist;
            }
        }
        if (plant == 6 ? bestDist < TH_HIGH : bestDist <= TH_HIGH) {
            if ((bestLevel == bestLevel2 || plant == 3) && bestDist > nnratio * bestDist2) continue;
            mp[bestIdx] = p;
            key_point[bestIdx] = p;
            pt_match[p] = bestIdx;
            nmatches++;
        }
    }
    return nmatches;
}

int oracle_search_by_projection_last(const oracle_keypoint* kp, const uint8_t* desc, const float* uright,
                                     const uint8_t* occupied, int n, const float* bounds, const float* scale,
                                     const float* cam, int mono, const oracle_last_point* pts,
                                     const uint8_t* pdesc, int npts, float th, int check_orientation,
                                     int32_t* pt_match, int32_t* key_point)
{
    const int plant = oracle_planted(4);
    int nmatches = 0;
    // Angles are cv::KeyPoint angles in [0, 360]; others would put a rotation
    // bin outside the histogram (the reference asserts, :1836).
    if (check_orientation) {
        for (int i = 0; i < n; i++)
            if (!(kp[i].angle >= 0.f && kp[i].angle <= 360.f)) return -1;
        for (int p = 0; p < npts; p++)
            if ((pts[p].flags & 1) && !(pts[p].angle >= 0.f && pts[p].angle <= 360.f)) return -1;
    }
    std::vector<int> rotHist[kHisto];
    const float factor = kHisto / 360.0f;
    const float fx = cam[0], fy = cam[1], cx = cam[2], cy = cam[3], mbf = cam[4], mb = cam[5], tlc_z = cam[6];
    const bool bForward = tlc_z > mb && !mono;
    const bool bBackward = -tlc_z > mb && !mono;
    const Grid g = build_grid(kp, n, bounds);
    std::vector<int> mp(n, -1);
    for (int i = 0; i < n; i++)
        if (occupied[i]) mp[i] = -2;
    auto has_obs = [&](int i) {
        return mp[i] == -2 || (mp[i] >= 0 && ((pts[mp[i]].flags & 4) || plant == 9));   // planted 9
    };
    for (int i = 0; i < n; i++) key_point[i] = -1;
    for (int p = 0; p < npts; p++) {
        pt_match[p] = -1;
        const oracle_last_point& P = pts[p];
        if (!(P.flags & 1)) continue;
        const float xc = P.xc, yc = P.yc;
        const float invzc = 1.0 / P.zc;
        if (invzc < 0) continue;
        const float u = fx * xc * invzc + cx;
        const float v = fy * yc * invzc + cy;
        if (u < bounds[0] || u > bounds[1]) continue;
        if (v < bounds[2] || v > bounds[3]) continue;
        const int nLastOctave = P.octave;
        const float radius = th * scale[nLastOctave];
        std::vector<int> vIndices2;
        if (bForward)
            vIndices2 = features_in_area(g, kp, u, v, radius, nLastOctave, -1);
        else if (bBackward)
            vIndices2 = plant == 7 ? features_in_area(g, kp, u, v, radius, nLastOctave - 1, nLastOctave)
                                   : features_in_area(g, kp, u, v, radius, 0, nLastOctave);
        else
            vIndices2 = features_in_area(g, kp, u, v, radius, nLastOctave - 1, nLastOctave + 1);
        if (vIndices2.empty()) continue;
        const uint8_t* dMP = pdesc + (size_t)p * 32;
        int bestDist = 256, bestIdx2 = -1;
        for (int i2 : vIndices2) {
            if (mp[i2] != -1 && plant != 4)
                if (has_obs(i2)) continue;
            if (plant == 4 && mp[i2] == -2) continue;
            if (uright[i2] > 0) {
                const float ur = u - mbf * invzc;
                const float er = std::fabs(ur - uright[i2]);
                if (plant == 5 ? er >= radius : er > radius) continue;
            }
            const int dist = forb_distance(dMP, desc + (size_t)i2 * 32);
            if (dist < bestDist) {
                bestDist = dist;
                bestIdx2 = i2;
            }
        }
        if (plant == 6 ? bestDist < TH_HIGH : bestDist <= TH_HIGH) {
            mp[bestIdx2] = p;
            key_point[bestIdx2] = p;
            pt_match[p] = bestIdx2;
            nmatches++;
            if (check_orientation) {
                float rot = P.angle - kp[bestIdx2].angle;
                if (rot < 0.0) rot += 360.0f;
                int bin = plant == 8 ? (int)std::nearbyint(rot * factor) : (int)std::round(rot * factor);
                if (bin == kHisto) bin = 0;
                rotHist[bin].push_back(bestIdx2);
            }
        }
    }
    if (check_orientation) {
        int ind1 = -1, ind2 = -1, ind3 = -1;
        three_maxima(rotHist, kHisto, ind1, ind2, ind3);
        for (int i = 0; i < kHisto; i++) {
            if (i == ind1 || i == ind2 || i == ind3) continue;
            for (int i2 : rotHist[i]) {
                key_point[i2] = -2;
                nmatches--;
            }
        }
    }
    return nmatches;
}

}  // extern "C"

