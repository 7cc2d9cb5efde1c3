// Features of the learned pruner and the labels it is trained on.
//
// A (tree T, dimension m) pair is the unit the level-synchronous driver would
// refine next: the neighbour counts of T's points in S ∪ {m}, the partition of
// T by the cell groups of m, and every restricted tree below.  Its features are
// what the method has in hand before those counts: T's size and depth, the
// dimension bitmask, how T's points spread over dimension m (value histogram,
// cell groups) and their counts in S.  All accumulations are integer, so the
// features, and with them the pruner's decisions, do not depend on the order
// in which blocks finish.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "inscy_device.h"
#include "inscy_internal.h"

#pragma clang fp contract(off)

namespace inscy {

namespace {

constexpr int kMaxCells = 256;  // cellC is uint8

// per pair: value histogram and cell histogram of T's points in dimension m,
// how many of them have a count in S at the cap and at least halfway from the
// weak threshold to the cap, and the sum of their counts
__global__ void k_pair_acc(const FeatPair *__restrict__ pairs, const int *__restrict__ blk_off, int np,
                           const float *__restrict__ Xt, const uint8_t *__restrict__ cellC, int n, int cells,
                           int weak_thr, int *__restrict__ acc, unsigned long long *__restrict__ accs) {
    __shared__ int h[kHistBins + kMaxCells + 2];
    __shared__ unsigned long long hs;
    const int j = seg_of(blk_off, np, blockIdx.x);
    const FeatPair P = pairs[j];
    const int stride = kHistBins + cells + 2;
    for (int e = threadIdx.x; e < stride; e += blockDim.x) h[e] = 0;
    if (threadIdx.x == 0) hs = 0ull;
    __syncthreads();
    const int i = (blockIdx.x - blk_off[j]) * blockDim.x + threadIdx.x;
    unsigned long long s = 0ull;
    if (i < P.nW) {
        const int p = P.W[i];
        const size_t at = (size_t) P.m * n + p;
        const float x = Xt[at];
        const int bin = min(kHistBins - 1, max(0, (int) (x * (float) kHistBins)));
        atomicAdd(&h[bin], 1);
        atomicAdd(&h[kHistBins + cellC[at]], 1);
        if (P.cntS) {
            const int c = min(max(P.cntS[p], 0), P.cap);
            if (c >= P.cap) atomicAdd(&h[kHistBins + cells], 1);
            if (2 * c >= weak_thr + P.cap) atomicAdd(&h[kHistBins + cells + 1], 1);
            s = (unsigned long long) c;
        }
    }
    // wave sum, then one LDS atomic per wave
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if (__lane_id() == 0 && s) atomicAdd(&hs, s);
    __syncthreads();
    int *out = acc + (size_t) j * stride;
    for (int e = threadIdx.x; e < stride; e += blockDim.x)
        if (h[e]) atomicAdd(&out[e], h[e]);
    if (threadIdx.x == 0 && hs) atomicAdd(&accs[j], hs);
}

__global__ void k_pair_finalize(const FeatPair *__restrict__ pairs, int np, int n, int d, int cells, int min_size,
                                const int *__restrict__ acc, const unsigned long long *__restrict__ accs,
                                float *__restrict__ feat) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= np) return;
    const FeatPair P = pairs[j];
    const int stride = kHistBins + cells + 2;
    const int *a = acc + (size_t) j * stride;
    float *f = feat + (size_t) j * kMlpIn;
    const float nW = (float) max(P.nW, 1);
    f[0] = (float) P.nW / (float) n;
    f[1] = log2f(1.0f + (float) P.nW) / 20.0f;
    f[2] = fminf(1.0f, (float) min_size / nW);
    f[3] = (float) P.k / 16.0f;
    f[4] = (float) (P.k + 1) / (float) d;
    f[5] = (float) P.m / (float) d;
    f[6] = (float) (d - P.first) / (float) d;
    f[7] = (float) P.nW / (float) max(P.pre, 1);
    // value histogram, largest bin first
    float hv[kHistBins];
    for (int b = 0; b < kHistBins; ++b) hv[b] = (float) a[b] / nW;
    for (int b = 1; b < kHistBins; ++b) {
        const float v = hv[b];
        int c = b;
        while (c > 0 && hv[c - 1] < v) {
            hv[c] = hv[c - 1];
            --c;
        }
        hv[c] = v;
    }
    for (int b = 0; b < kHistBins; ++b) f[8 + b] = hv[b];
    // cell groups of dimension m (group ids ascend with the cells)
    int big = 0, nbig = 0, mass_big = 0;
    {
        int cur = -1, sum = 0;
        for (int c = 0; c <= cells; ++c) {
            const int g = c < cells ? P.gid[c] : -2;
            if (g != cur) {
                if (cur >= 0) {
                    big = max(big, sum);
                    if (sum >= min_size) {
                        ++nbig;
                        mass_big += sum;
                    }
                }
                cur = g;
                sum = 0;
            }
            if (c < cells) sum += a[kHistBins + c];
        }
    }
    f[24] = (float) big / nW;
    f[25] = (float) P.ngroups / (float) cells;
    f[26] = (float) mass_big / nW;
    f[27] = (float) nbig / (float) cells;
    const bool hasS = P.cntS != nullptr;
    f[28] = hasS ? (float) a[kHistBins + cells] / nW : 0.0f;
    f[29] = hasS ? (float) a[kHistBins + cells + 1] / nW : 0.0f;
    f[30] = hasS ? (float) ((double) accs[j] / ((double) nW * (double) max(P.cap, 1))) : 0.0f;
    f[31] = hasS ? 1.0f : 0.0f;
    for (int b = 0; b < 32; ++b) f[32 + b] = 0.0f;
    for (dmask_t mk = P.mask; mk; mk &= mk - 1) f[32 + dfirst(mk) % 32] = 1.0f;
}

__global__ void k_tree_hits(const TreeHit *__restrict__ trees, const int *__restrict__ blk_off, int nt,
                            int *__restrict__ hit) {
    const int t = seg_of(blk_off, nt, blockIdx.x);
    const TreeHit T = trees[t];
    const int i = (blockIdx.x - blk_off[t]) * blockDim.x + threadIdx.x;
    bool h = false;
    if (i < T.nW) {
        const int c = T.comp[T.W[i]];
        h = c >= 0 && T.inres[c];
    }
    if (__ballot(h) && __lane_id() == 0) atomicOr(&hit[T.out], 1);
}

}  // namespace

hipError_t launch_pair_features(const FeatPair *pairs, const int *blk_off, int npairs, int nblk, const float *Xt,
                                const uint8_t *cellC, int n, int d, int cells, int min_size, int weak_thr, int *acc,
                                unsigned long long *accs, float *feat, hipStream_t st) {
    if (npairs <= 0) return hipSuccess;
    if (cells > kMaxCells) return hipErrorInvalidValue;
    if (nblk > 0)
        hipLaunchKernelGGL(k_pair_acc, dim3(nblk), dim3(kBlock), 0, st, pairs, blk_off, npairs, Xt, cellC, n, cells,
                           weak_thr, acc, accs);
    hipLaunchKernelGGL(k_pair_finalize, dim3(blocks_for(npairs)), dim3(kBlock), 0, st, pairs, npairs, n, d, cells,
                       min_size, acc, accs, feat);
    return hipGetLastError();
}

hipError_t launch_tree_hits(const TreeHit *trees, const int *blk_off, int ntrees, int nblk, int *hit, hipStream_t st) {
    if (ntrees <= 0 || nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tree_hits, dim3(nblk), dim3(kBlock), 0, st, trees, blk_off, ntrees, hit);
    return hipGetLastError();
}

}  // namespace inscy
