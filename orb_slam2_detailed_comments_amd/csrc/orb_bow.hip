// orb_bow.hip -- gfx950 kernels for the DBoW2 vocabulary transform
// (Thirdparty/DBoW2/DBoW2/TemplatedVocabulary.h), as ORB-SLAM2 calls it from
// Frame::ComputeBoW (src/Frame.cc:704-718):
//   voc_transform_kernel  transform(feature, word_id, weight, nid, levelsup)
//                         (TemplatedVocabulary.h:1240-1290) for every row
//   voc_bow_kernel        transform(features, BowVector, FeatureVector,
//                         levelsup) (:1136-1206): per image, the word-sorted
//                         BowVector with the reference's weighting and
//                         normalisation, and the node-sorted FeatureVector
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "orb_types.h"

#define VOC_LANES 16                 // lanes per descriptor: one per child
#define BOW_CAP 4096                 // descriptors per image (LDS sort)

// Descend from the root: at each level the 16 lanes of a group take one
// child each (more than 16 children loop), key = distance << 22 | child
// index (the host admits < 2^22 children a node), and the group minimum is
// the reference's strict-< first minimum.
// Each lane loads its child's descriptor and its child's (first, count) in
// the same round, so the chosen child's children are known without another
// dependent load.
__global__ __launch_bounds__(256) void voc_transform_kernel(VocDev V, const uint8_t* __restrict__ desc, int n,
                                                            int levelsup, int* __restrict__ word_out,
                                                            double* __restrict__ weight_out, int* __restrict__ nid_out)
{
    const int sub = threadIdx.x & (VOC_LANES - 1);
    const int row = blockIdx.x * (256 / VOC_LANES) + (threadIdx.x / VOC_LANES);
    const bool live = row < n;
    const int rr = live ? row : 0;
    const uint4 q0 = *(const uint4*)(desc + (size_t)rr * 32), q1 = *(const uint4*)(desc + (size_t)rr * 32 + 16);
    const int nid_level = V.L - levelsup;
    int node = 0, nid = 0;
    int2 meta = V.meta[0];
    for (int level = 1; level <= V.max_depth && meta.y > 0; level++) {
        unsigned best = 0xFFFFFFFFu;
        int2 bmeta = make_int2(0, 0);
        for (int j = sub; j < meta.y; j += VOC_LANES) {
            const int c = meta.x + j;
            const uint4 d0 = V.desc[2 * c], d1 = V.desc[2 * c + 1];
            const int2 cm = V.meta[c];
            const unsigned d = __popc(q0.x ^ d0.x) + __popc(q0.y ^ d0.y) + __popc(q0.z ^ d0.z) +
                               __popc(q0.w ^ d0.w) + __popc(q1.x ^ d1.x) + __popc(q1.y ^ d1.y) +
                               __popc(q1.z ^ d1.z) + __popc(q1.w ^ d1.w);
            const unsigned key = (d << 22) | (unsigned)j;
            if (key < best) {
                best = key;
                bmeta = cm;
            }
        }
#pragma unroll
        for (int o = VOC_LANES / 2; o > 0; o >>= 1) {
            const unsigned ob = __shfl_xor(best, o, VOC_LANES);
            const int ox = __shfl_xor(bmeta.x, o, VOC_LANES), oy = __shfl_xor(bmeta.y, o, VOC_LANES);
            if (ob < best) {
                best = ob;
                bmeta = make_int2(ox, oy);
            }
        }
        node = meta.x + (int)(best & 0x3FFFFFu);
        meta = bmeta;
        if (level == nid_level) nid = node;
    }
    if (live && sub == 0) {
        word_out[row] = V.word_id[node];
        weight_out[row] = V.weight[node];
        if (nid_out) nid_out[row] = nid_level <= 0 ? 0 : V.orig_id[nid];
    }
}

__device__ __forceinline__ void bitonic_sort_u64(unsigned long long* v, int npow)
{
    for (int k = 2; k <= npow; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int a = threadIdx.x; a < npow; a += blockDim.x) {
                const int b = a ^ j;
                if (b > a) {
                    const unsigned long long va = v[a], vb = v[b];
                    const bool up = (a & k) == 0;
                    if (up ? (va > vb) : (va < vb)) {
                        v[a] = vb;
                        v[b] = va;
                    }
                }
            }
            __syncthreads();
        }
}

// One workgroup per image (1024 threads, thread t owns sorted positions
// 4t..4t+3).  Rows with weight > 0 only (the reference's "not stopped" test).
// BowVector: word-sorted; under TF / TF_IDF a word's value is its rows'
// weights added in feature order (addWeight), under IDF / BINARY the weight
// of its first row (addIfNotExist); TF / TF_IDF without normalisation then
// divide by the number of words, and normalising scorings divide by the L1 or
// L2 norm summed in word order (BowVector::normalize).  FeatureVector:
// node-sorted, feature indices ascending within a node (addFeature in index
// order).
__device__ __forceinline__ int block_excl_scan_1024(int x, int* sc)
{
    const int t = threadIdx.x;
    sc[t] = x;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int y = t >= o ? sc[t - o] : 0;
        __syncthreads();
        sc[t] += y;
        __syncthreads();
    }
    const int incl = sc[t];
    __syncthreads();
    return incl - x;
}

__global__ __launch_bounds__(1024) void voc_bow_kernel(VocDev V, const BowJob* __restrict__ jobs,
                                                       const int* __restrict__ word, const double* __restrict__ wgt,
                                                       const int* __restrict__ nidv, int* __restrict__ bow_word,
                                                       double* __restrict__ bow_value, int* __restrict__ n_bow,
                                                       int* __restrict__ fv_node, int* __restrict__ fv_feat,
                                                       int* __restrict__ n_fv)
{
    __shared__ unsigned long long key[BOW_CAP];   // sort keys; later the word values (as double)
    __shared__ int wordv[BOW_CAP];
    __shared__ int sc[1024];
    __shared__ int s_cnt;
    __shared__ double s_norm;
    const BowJob J = jobs[blockIdx.x];
    const int n = min(J.n, BOW_CAP);
    const int tid = threadIdx.x;
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    for (int i = tid; i < n; i += 1024)
        if (wgt[J.off + i] > 0) key[atomicAdd(&s_cnt, 1)] = ((unsigned long long)(unsigned)word[J.off + i] << 32) | i;
    __syncthreads();
    const int m = s_cnt;
    int npow = 1;
    while (npow < m) npow <<= 1;
    for (int i = m + tid; i < npow; i += 1024) key[i] = ~0ull;
    __syncthreads();
    bitonic_sort_u64(key, npow);
    const bool tf = V.weighting == 0 || V.weighting == 1;
    const bool must = V.scoring != 5;
    const bool l2 = V.scoring == 1;
    // heads of word runs among this thread's four positions, and their values
    int hw[4];
    double hv[4];
    int nh = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int i = 4 * tid + e;
        hw[e] = -1;
        hv[e] = 0.0;
        if (i >= m) continue;
        const unsigned w = (unsigned)(key[i] >> 32);
        if (i > 0 && (unsigned)(key[i - 1] >> 32) == w) continue;
        double v = wgt[J.off + (int)(unsigned)key[i]];
        if (tf)
            for (int r = i + 1; r < m && (unsigned)(key[r] >> 32) == w; r++) v += wgt[J.off + (int)(unsigned)key[r]];
        hw[e] = (int)w;
        hv[e] = v;
        nh++;
    }
    const int rank0 = block_excl_scan_1024(nh, sc);
    __syncthreads();                              // every key read above is done
    double* val = reinterpret_cast<double*>(key);
    int r = rank0;
#pragma unroll
    for (int e = 0; e < 4; e++)
        if (hw[e] >= 0) {
            wordv[r] = hw[e];
            val[r] = hv[e];
            r++;
        }
    if (tid == 1023) sc[0] = r;                   // total heads (last thread's end rank)
    __syncthreads();
    const int nb = sc[0];
    if (tid == 0) {
        double norm = 0.0;
        if (must) {
            if (!l2) {
                for (int i = 0; i < nb; i++) norm += fabs(val[i]);
            } else {
                for (int i = 0; i < nb; i++) norm += val[i] * val[i];
                norm = sqrt(norm);
            }
        }
        s_norm = norm;
        n_bow[blockIdx.x] = nb;
    }
    __syncthreads();
    const double norm = s_norm;
    for (int i = tid; i < nb; i += 1024) {
        double v = val[i];
        if (tf && !must) v = v / (double)nb;
        if (must && norm > 0.0) v = v / norm;
        bow_word[J.off + i] = wordv[i];
        bow_value[J.off + i] = v;
    }
    __syncthreads();
    // FeatureVector: (node << 32 | feature) over the same kept rows
    if (tid == 0) s_cnt = 0;
    __syncthreads();
    for (int i = tid; i < n; i += 1024)
        if (wgt[J.off + i] > 0) key[atomicAdd(&s_cnt, 1)] = ((unsigned long long)(unsigned)nidv[J.off + i] << 32) | i;
    __syncthreads();
    for (int i = m + tid; i < npow; i += 1024) key[i] = ~0ull;
    __syncthreads();
    bitonic_sort_u64(key, npow);
    for (int i = tid; i < m; i += 1024) {
        fv_node[J.off + i] = (int)(unsigned)(key[i] >> 32);
        fv_feat[J.off + i] = (int)(unsigned)key[i];
    }
    if (tid == 0) n_fv[blockIdx.x] = m;
}

extern "C" hipError_t orb_launch_voc_transform(const VocDev* V, const uint8_t* desc, int n, int levelsup,
                                               int* word, double* weight, int* nid, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int per_block = 256 / VOC_LANES;
    voc_transform_kernel<<<(n + per_block - 1) / per_block, 256, 0, st>>>(*V, desc, n, levelsup, word, weight, nid);
    return hipGetLastError();
}

extern "C" hipError_t orb_launch_voc_bow(const VocDev* V, const BowJob* jobs, int n_images, const int* word,
                                         const double* weight, const int* nid, int* bow_word, double* bow_value,
                                         int* n_bow, int* fv_node, int* fv_feat, int* n_fv, hipStream_t st)
{
    if (n_images <= 0) return hipSuccess;
    voc_bow_kernel<<<n_images, 1024, 0, st>>>(*V, jobs, word, weight, nid, bow_word, bow_value, n_bow, fv_node,
                                              fv_feat, n_fv);
    return hipGetLastError();
}
