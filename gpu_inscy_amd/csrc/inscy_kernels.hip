// HIP kernels of the MI355X INSCY engine (gfx950, wave64).
//
// Everything here is integer/byte streaming work plus one FP32 distance loop;
// nothing is GEMM-shaped (the contraction depth is |S| <= a few dims), so the
// bound for every kernel is HBM/L2 bandwidth or VALU issue, not MFMA.
//
// Floating point: the reference compares fl(sqrt(sum_k fl(fl(p_k - q_k)^2))) <= eps
// with separate multiplies and adds in subspace order (ClusteringCpu.cu:40-81).
// This file is compiled with -ffp-contract=off and the pragma below so that no
// multiply-add is fused, and the sqrt is replaced by the exact equivalent
// `sum <= s_star` (s_star = largest float whose sqrtf is <= eps).

#include <climits>
#include <cmath>
#include <cstring>

#include "inscy_internal.h"

#pragma clang fp contract(off)

namespace inscy {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int ld_acq(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_rel(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-aggregated slot allocation for order-free stream compaction: one
// atomic per wave instead of one per kept element.  Must be reached by every
// active lane of the wave.
__device__ __forceinline__ int wave_slot(bool keep, int *counter) {
    unsigned long long m = __ballot(keep);
    if (m == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll((long long) m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    return keep ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// ------------------------------------------------------------ discretisation
// get_cell_no (ScyTreeNode.cu:100-102): min((int)(x / (1/c)), c-1); the
// s-connection test of construct_s_connection (:34-49): x >= (c+1)*(1/c) - eps.
__global__ void k_discretise(const float *__restrict__ X, int n, int d, int cells, float cs, int ecells,
                             float ecs, float eps, float *__restrict__ Xt, uint8_t *__restrict__ cellC,
                             uint16_t *__restrict__ cellE, uint64_t *__restrict__ nearm, int *__restrict__ bad) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint64_t nm = 0;
    bool out_of_range = false;
    for (int k = 0; k < d; ++k) {
        float x = X[(size_t) p * d + k];
        if (!(x >= 0.0f && x <= 1.0f)) out_of_range = true;
        x = fminf(fmaxf(x, 0.0f), 1.0f);  // only reached with valid data when bad stays 0
        if (Xt) Xt[(size_t) k * n + p] = x;
        int c = min((int) __fdiv_rn(x, cs), cells - 1);
        int e = min((int) __fdiv_rn(x, ecs), ecells - 1);
        cellC[(size_t) k * n + p] = (uint8_t) c;
        cellE[(size_t) k * n + p] = (uint16_t) e;
        float upper = __fmul_rn((float) (c + 1), cs);
        float bound = __fsub_rn(upper, eps);
        if (x >= bound) nm |= 1ull << k;
    }
    nearm[p] = nm;
    if (out_of_range) atomicOr(bad, 1);
}

// ------------------------------------------------------------ per-dimension eps-grid buckets
__global__ void k_hist1(const uint16_t *__restrict__ cellE, int n, int d, int ecells, int *__restrict__ hist) {
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t) n * d) return;
    int k = (int) (i / n);
    atomicAdd(&hist[(size_t) k * (ecells + 1) + cellE[i]], 1);
}

// one block per dimension: exclusive scan of ecells counts into off1[k][0..ecells]
__global__ void k_scan1(int *__restrict__ hist, int ecells, int *__restrict__ off1) {
    int k = blockIdx.x;
    int *h = hist + (size_t) k * (ecells + 1);
    int *o = off1 + (size_t) k * (ecells + 1);
    __shared__ int part[kBlock];
    int per = (ecells + kBlock - 1) / kBlock;
    int beg = threadIdx.x * per, end = min(beg + per, ecells);
    int s = 0;
    for (int i = beg; i < end; ++i) s += h[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < kBlock; ++t) {
            int v = part[t];
            part[t] = acc;
            acc += v;
        }
        o[ecells] = acc;
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int i = beg; i < end; ++i) {
        int v = h[i];
        o[i] = acc;
        h[i] = acc;  // becomes the scatter cursor
        acc += v;
    }
}

__global__ void k_scatter1(const uint16_t *__restrict__ cellE, int n, int d, int ecells, int *__restrict__ cursor,
                           int *__restrict__ sid1) {
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t) n * d) return;
    int k = (int) (i / n);
    int p = (int) (i - (size_t) k * n);
    int pos = atomicAdd(&cursor[(size_t) k * (ecells + 1) + cellE[i]], 1);
    sid1[(size_t) k * n + pos] = p;
}

// ------------------------------------------------------------ restriction
// mergeWithNeighbors decision input (ScyTreeNode.cu:837-851 with the flag of
// restrict_node :262-267): cell c of dimension m is s-connected iff a path
// branching at m starts from a point in cell c.
__global__ void k_flags(const int *__restrict__ spx, const uint64_t *__restrict__ spm, const int *__restrict__ nsp,
                        const uint8_t *__restrict__ cellC_m, int m, int *__restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsp) return;
    if ((spm[i] >> m) & 1ull) flags[cellC_m[spx[i]]] = 1;
}

// Points and s-connection paths of the region {cell_m in [lo, hi]}; paths keep
// only branch dimensions above m (the only ones any descendant can use).
__global__ void k_filter(const int *__restrict__ P, const int *__restrict__ nP, const int *__restrict__ spx,
                         const uint64_t *__restrict__ spm, const int *__restrict__ nsp,
                         const uint8_t *__restrict__ cellC_m, int m, int lo, int hi, int *__restrict__ outP,
                         int *__restrict__ outx, uint64_t *__restrict__ outm, int *__restrict__ counters,
                         int nP_blocks) {
    if ((int) blockIdx.x < nP_blocks) {
        int i = blockIdx.x * blockDim.x + threadIdx.x;
        bool keep = false;
        int p = 0;
        if (i < *nP) {
            p = P[i];
            int c = cellC_m[p];
            keep = c >= lo && c <= hi;
        }
        int slot = wave_slot(keep, &counters[0]);
        if (keep) outP[slot] = p;
    } else {
        int i = (blockIdx.x - nP_blocks) * blockDim.x + threadIdx.x;
        bool keep = false;
        int x = 0;
        uint64_t mk = 0;
        if (i < *nsp) {
            x = spx[i];
            mk = m >= 63 ? 0ull : (spm[i] & (~0ull << (m + 1)));
            int c = cellC_m[x];
            keep = mk != 0 && c >= lo && c <= hi;
        }
        int slot = wave_slot(keep, &counters[1]);
        if (keep) {
            outx[slot] = x;
            outm[slot] = mk;
        }
    }
}

// ------------------------------------------------------------ epsilon neighbourhood
// Candidate rule of get_possible_neighbors_from (ScyTreeNode.cu:896-946) on
// the ceil(1/eps)-grid: |cell_e(q) - cell_e(p)| <= 1 in every dimension of S.
// The scan runs over the three eps-cells of the dimension of S whose slab
// around the query holds the fewest points; the other dimensions are checked
// per candidate.  Query coordinates and cells sit in LDS ([k][block] layout,
// conflict-free across the wave).
template <class Visit>
__device__ __forceinline__ long long visit_neighbours(const DevData &D, const SubspaceArg &S, int q,
                                                      const float *__restrict__ qx, const int *__restrict__ qe,
                                                      float s_star, Visit visit) {
    const int n = D.n;
    const int stride = (D.ecells + 1);
    int best = 0, a = 0, b = 0;
    int bsize = INT_MAX;
    for (int t = 0; t < S.k; ++t) {
        int s = S.dims[t];
        int e = qe[t * kBlock];
        int lo = max(e - 1, 0), hi = min(e + 1, D.ecells - 1);
        int aa = D.off1[(size_t) s * stride + lo], bb = D.off1[(size_t) s * stride + hi + 1];
        if (bb - aa < bsize) {
            bsize = bb - aa;
            best = s;
            a = aa;
            b = bb;
        }
    }
    const int *ids = D.sid1 + (size_t) best * n;
    for (int j = a; j < b; ++j) {
        int o = ids[j];
        if (o == q) continue;
        bool ok = true;
        float acc = 0.0f;
        for (int t = 0; t < S.k && ok; ++t) {
            int s = S.dims[t];
            int de = (int) D.cellE[(size_t) s * n + o] - qe[t * kBlock];
            float diff = qx[t * kBlock] - D.Xt[(size_t) s * n + o];
            float sq = diff * diff;
            acc = acc + sq;
            ok = de >= -1 && de <= 1 && acc <= s_star;
        }
        if (ok) visit(o);
    }
    return (long long) (b - a);
}

__device__ __forceinline__ void stage_query(const DevData &D, const SubspaceArg &S, int q, float *qx, int *qe) {
    for (int t = 0; t < S.k; ++t) {
        int s = S.dims[t];
        qx[t * kBlock] = D.Xt[(size_t) s * D.n + q];
        qe[t * kBlock] = D.cellE[(size_t) s * D.n + q];
    }
}

__global__ __launch_bounds__(kBlock) void k_counts(DevData D, SubspaceArg S, const int *__restrict__ queries,
                                                   const int *__restrict__ nq, float s_star, int *__restrict__ cnt,
                                                   unsigned long long *__restrict__ cand) {
    extern __shared__ char smem[];
    float *sx = reinterpret_cast<float *>(smem);
    int *se = reinterpret_cast<int *>(smem + sizeof(float) * kBlock * S.k);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    long long seen = 0, done = 0;
    if (i < *nq) {
        int q = queries[i];
        if (cnt[q] < 0) {
            float *qx = sx + threadIdx.x;
            int *qe = se + threadIdx.x;
            stage_query(D, S, q, qx, qe);
            int c = 0;
            seen = visit_neighbours(D, S, q, qx, qe, s_star, [&](int) { ++c; });
            cnt[q] = c;
            done = 1;
        }
    }
    if (cand) {  // cand[0]: candidates examined, cand[1]: queries computed
        for (int off = 32; off > 0; off >>= 1) {
            seen += __shfl_down(seen, off);
            done += __shfl_down(done, off);
        }
        if (__lane_id() == 0 && done) {
            atomicAdd(&cand[0], (unsigned long long) seen);
            atomicAdd(&cand[1], (unsigned long long) done);
        }
    }
}

// weak-density prune (pruneRecursionAndRemove2, ScyTreeNode.cu:582-667): keep
// points with |N_S(p)| >= thr; stamp them with the tree id for O(1) membership.
__global__ void k_weak(const int *__restrict__ P, const int *__restrict__ nP, const int *__restrict__ cnt, int thr,
                       int *__restrict__ W, int *__restrict__ nW, int *__restrict__ stamp, int tid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    int p = 0;
    if (i < *nP) {
        p = P[i];
        keep = cnt[p] >= thr;
    }
    int slot = wave_slot(keep, nW);
    if (keep) {
        W[slot] = p;
        stamp[p] = tid;
    }
}

// ------------------------------------------------------------ s-connection paths after the prune
// A path (b, x) survives the prune of tree R iff some surviving point w shares
// x's cells on every unrestricted dimension <= b.  (propergate_count2,
// ScyTreeNode.cu:446-467, erases the s-connection whose same-cell regular
// sibling lost all its points; the regular node a path hangs from always keeps
// a point as long as that sibling does.)  Membership is answered by an
// open-addressing table of 64-bit entries (32 hash bits | representative point
// id); every hash match is confirmed by comparing the cells themselves, so the
// test is exact.

constexpr uint64_t kEmpty = ~0ull;

__device__ __forceinline__ uint64_t mix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

__device__ __forceinline__ bool same_prefix(const DevData &D, uint64_t smask, int a, int b, int upto) {
    if (b < 0 || b >= D.n) return false;
    for (int k = 0; k <= upto; ++k) {
        if ((smask >> k) & 1ull) continue;
        if (D.cellC[(size_t) k * D.n + a] != D.cellC[(size_t) k * D.n + b]) return false;
    }
    return true;
}

__global__ void k_path_bits(const int *__restrict__ spx, const uint64_t *__restrict__ spm, const int *__restrict__ nsp,
                            const int *__restrict__ stamp, int tid, unsigned long long *__restrict__ bits) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t m = 0;
    if (i < *nsp && stamp[spx[i]] != tid) m = spm[i];
    // wave OR-reduction, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off);
    if (__lane_id() == 0 && m) atomicOr(bits, (unsigned long long) m);
}

__global__ void k_path_insert(DevData D, uint64_t smask, const int *__restrict__ W, const int *__restrict__ nW,
                              const unsigned long long *__restrict__ bits, uint64_t *__restrict__ table,
                              uint64_t tmask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nW) return;
    uint64_t need = *bits;
    if (!need) return;
    int w = W[i];
    uint64_t h = 0x243F6A8885A308D3ull;
    for (int k = 0; k < D.d; ++k) {
        if ((smask >> k) & 1ull) continue;
        h = mix64(h + D.cellC[(size_t) k * D.n + w] + 1);
        if (!((need >> k) & 1ull)) continue;
        uint64_t e = mix64(h ^ (0x9E3779B97F4A7C15ull * (uint64_t) (k + 1)));
        uint64_t entry = (e & 0xFFFFFFFF00000000ull) | (uint32_t) w;
        uint64_t slot = e & tmask;
        while (true) {
            uint64_t cur = __hip_atomic_load(&table[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == kEmpty) {
                cur = atomicCAS((unsigned long long *) &table[slot], (unsigned long long) kEmpty,
                                (unsigned long long) entry);
                if (cur == kEmpty) break;
            }
            if ((cur >> 32) == (entry >> 32) && same_prefix(D, smask, w, (int) (uint32_t) cur, k)) break;
            slot = (slot + 1) & tmask;
        }
    }
}

__global__ void k_path_check(DevData D, uint64_t smask, const int *__restrict__ spx, const uint64_t *__restrict__ spm,
                             const int *__restrict__ nsp, const int *__restrict__ stamp, int tid,
                             const uint64_t *__restrict__ table, uint64_t tmask, int *__restrict__ outx,
                             uint64_t *__restrict__ outm, int *__restrict__ nout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int x = 0;
    uint64_t m = 0;
    if (i < *nsp) {
        x = spx[i];
        m = spm[i];
        if (stamp[x] != tid) {
            uint64_t h = 0x243F6A8885A308D3ull;
            uint64_t keep = 0;
            for (int k = 0; k < D.d; ++k) {
                if ((smask >> k) & 1ull) continue;
                h = mix64(h + D.cellC[(size_t) k * D.n + x] + 1);
                if (!((m >> k) & 1ull)) continue;
                uint64_t e = mix64(h ^ (0x9E3779B97F4A7C15ull * (uint64_t) (k + 1)));
                uint64_t slot = e & tmask;
                while (true) {
                    uint64_t cur = __hip_atomic_load(&table[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (cur == kEmpty) break;
                    if ((cur >> 32) == (e >> 32) && same_prefix(D, smask, x, (int) (uint32_t) cur, k)) {
                        keep |= 1ull << k;
                        break;
                    }
                    slot = (slot + 1) & tmask;
                }
            }
            m = keep;
        }
    }
    int slot = wave_slot(m != 0, nout);
    if (m != 0) {
        outx[slot] = x;
        outm[slot] = m;
    }
}

// ------------------------------------------------------------ dense components
__global__ void k_all_queries(const int *__restrict__ cnt, int n, int *__restrict__ q, int *__restrict__ nq) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool need = i < n && cnt[i] < 0;
    int slot = wave_slot(need, nq);
    if (need) q[slot] = i;
}

__device__ __forceinline__ int uf_find(int *par, int x) {
    int p = ld_acq(&par[x]);
    while (p != x) {
        int g = ld_acq(&par[p]);
        if (g != p) st_rel(&par[x], g);  // path halving: g is an ancestor of x
        x = p;
        p = g;
    }
    return x;
}

__device__ __forceinline__ void uf_union(int *par, int a, int b) {
    while (true) {
        a = uf_find(par, a);
        b = uf_find(par, b);
        if (a == b) return;
        if (a > b) {
            int t = a;
            a = b;
            b = t;
        }
        // hook the larger root under the smaller: roots stay component minima
        int old = atomicCAS(&par[b], b, a);
        if (old == b) return;
        b = old;
    }
}

__global__ void k_uf_init(const int *__restrict__ cnt, int n, int thr, int *__restrict__ par) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) par[i] = cnt[i] >= thr ? i : -1;
}

__global__ __launch_bounds__(kBlock) void k_uf_link(DevData D, SubspaceArg S, float s_star, int *__restrict__ par) {
    extern __shared__ char smem[];
    float *sx = reinterpret_cast<float *>(smem);
    int *se = reinterpret_cast<int *>(smem + sizeof(float) * kBlock * S.k);
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= D.n || ld_acq(&par[q]) < 0) return;
    float *qx = sx + threadIdx.x;
    int *qe = se + threadIdx.x;
    stage_query(D, S, q, qx, qe);
    visit_neighbours(D, S, q, qx, qe, s_star, [&](int o) {
        if (o > q && ld_acq(&par[o]) >= 0) uf_union(par, q, o);
    });
}

__global__ void k_uf_flatten(int *__restrict__ par, int n, int *__restrict__ comp, int *__restrict__ csize) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = -1;
    if (ld_acq(&par[i]) >= 0) {
        c = uf_find(par, i);
        atomicAdd(&csize[c], 1);
    }
    comp[i] = c;
}

// ------------------------------------------------------------ redundancy and join
// pruneRedundancy (ScyTreeNode.cu:669-734): per superspace H, the smallest
// result cluster of H that holds a point of the region.
__global__ void k_redundancy(const int *__restrict__ W, const int *__restrict__ nW, int const *const *compH,
                             int const *const *csizeH, uint8_t const *const *inresH, int *__restrict__ mn) {
    int h = blockIdx.y;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int best = INT_MAX;
    if (i < *nW) {
        int w = W[i];
        int c = compH[h][w];
        if (c >= 0 && inresH[h][c]) best = csizeH[h][c];
    }
    for (int off = 32; off > 0; off >>= 1) best = min(best, __shfl_down(best, off));
    if (__lane_id() == 0 && best != INT_MAX) atomicMin(&mn[h], best);
}

// cluster (ClusteringCpu.cu:317-389): the BFS from region point w labels w's
// dense component; here the component id is marked as labelled.
__global__ void k_seed(const int *__restrict__ W, const int *__restrict__ nW, const int *__restrict__ comp,
                       uint8_t *__restrict__ lab) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nW) return;
    int c = comp[W[i]];
    if (c >= 0) lab[c] = 1;
}

__global__ void k_collect(const uint8_t *__restrict__ lab, int n, const int *__restrict__ csize, int *__restrict__ out,
                          int *__restrict__ nout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = i < n && lab[i];
    int slot = wave_slot(keep, nout);
    if (keep) {
        out[2 * slot] = i;
        out[2 * slot + 1] = csize[i];
    }
}

// join (util.cu:655-743): a cluster K of S is dropped when one of its points
// lies in a result cluster of a superspace H that is larger than |K| * r.
__global__ void k_join_check(const int *__restrict__ compS, const uint8_t *__restrict__ keep,
                             const int *__restrict__ csizeS, int n, float r, int const *const *compH,
                             int const *const *csizeH, uint8_t const *const *inresH, uint8_t *__restrict__ drop) {
    int h = blockIdx.y;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = compS[i];
    if (k < 0 || !keep[k]) return;
    int kh = compH[h][i];
    if (kh < 0 || !inresH[h][kh]) return;
    if ((float) csizeS[k] * r < (float) csizeH[h][kh]) drop[k] = 1;
}

__global__ void k_labels(const int *__restrict__ comp, const uint8_t *__restrict__ inres, int n, int *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = comp[i];
    out[i] = (c >= 0 && inres[c]) ? c : -1;
}

template <class T>
__global__ void k_fill(T *p, T v, size_t count) {
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

__global__ void k_set_flags(uint8_t *flags, const int *ids, int count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) flags[ids[i]] = 1;
}

__global__ void k_iota(int *__restrict__ p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Root s-connection paths: every point with at least one near-edge dimension.
__global__ void k_root_paths(const uint64_t *__restrict__ nearm, int n, int *__restrict__ outx,
                             uint64_t *__restrict__ outm, int *__restrict__ nout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t m = i < n ? nearm[i] : 0ull;
    int slot = wave_slot(m != 0, nout);
    if (m != 0) {
        outx[slot] = i;
        outm[slot] = m;
    }
}

inline unsigned blocks_for(size_t count) { return (unsigned) ((count + kBlock - 1) / kBlock); }

}  // namespace

// ------------------------------------------------------------ launchers
float sqrt_threshold(float eps) {
    // bisection on the ordered bit patterns of non-negative floats
    uint32_t lo = 0, hi = 0x7f800000u;  // [0, +inf)
    auto sq = [](uint32_t b) {
        float f;
        std::memcpy(&f, &b, 4);
        return sqrtf(f);
    };
    if (!(sq(lo) <= eps)) return -1.0f;
    while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (sq(mid) <= eps)
            lo = mid;
        else
            hi = mid;
    }
    float f;
    std::memcpy(&f, &lo, 4);
    return f;
}

hipError_t launch_discretise(const float *X, int n, int d, int cells, float eps, float *Xt, uint8_t *cellC,
                             uint16_t *cellE, uint64_t *nearm, int *bad, hipStream_t st) {
    int ecells = (int) ceil(1.0 / (double) eps);
    float cs = 1.0f / (float) cells;
    float ecs = 1.0f / (float) ecells;
    hipLaunchKernelGGL(k_discretise, dim3(blocks_for(n)), dim3(kBlock), 0, st, X, n, d, cells, cs, ecells, ecs, eps,
                       Xt, cellC, cellE, nearm, bad);
    return hipGetLastError();
}

hipError_t launch_build_index(const uint16_t *cellE, int n, int d, int ecells, int *hist, int *off1, int *sid1,
                              hipStream_t st) {
    size_t nd = (size_t) n * d;
    hipLaunchKernelGGL(k_fill<int>, dim3(blocks_for((size_t) d * (ecells + 1))), dim3(kBlock), 0, st, hist, 0,
                       (size_t) d * (ecells + 1));
    hipLaunchKernelGGL(k_hist1, dim3(blocks_for(nd)), dim3(kBlock), 0, st, cellE, n, d, ecells, hist);
    hipLaunchKernelGGL(k_scan1, dim3(d), dim3(kBlock), 0, st, hist, ecells, off1);
    hipLaunchKernelGGL(k_scatter1, dim3(blocks_for(nd)), dim3(kBlock), 0, st, cellE, n, d, ecells, hist, sid1);
    return hipGetLastError();
}

hipError_t launch_flags(const int *spx, const uint64_t *spm, const int *nsp, int nsp_max, const uint8_t *cellC_m,
                        int m, int *flags, hipStream_t st) {
    if (nsp_max <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_flags, dim3(blocks_for(nsp_max)), dim3(kBlock), 0, st, spx, spm, nsp, cellC_m, m, flags);
    return hipGetLastError();
}

hipError_t launch_filter(const int *P, const int *nP, int nP_max, const int *spx, const uint64_t *spm,
                         const int *nsp, int nsp_max, const uint8_t *cellC_m, int m, int lo, int hi, int *outP,
                         int *outx, uint64_t *outm, int *counters, hipStream_t st) {
    int bp = (int) blocks_for(nP_max), bs = (int) blocks_for(nsp_max);
    if (bp + bs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_filter, dim3(bp + bs), dim3(kBlock), 0, st, P, nP, spx, spm, nsp, cellC_m, m, lo, hi, outP,
                       outx, outm, counters, bp);
    return hipGetLastError();
}

hipError_t launch_counts(const DevData &D, const SubspaceArg &S, const int *queries, const int *nq, int nq_max,
                         float s_star, int *cnt, unsigned long long *cand, hipStream_t st) {
    if (nq_max <= 0) return hipSuccess;
    size_t lds = (sizeof(float) + sizeof(int)) * kBlock * S.k;
    hipLaunchKernelGGL(k_counts, dim3(blocks_for(nq_max)), dim3(kBlock), lds, st, D, S, queries, nq, s_star, cnt, cand);
    return hipGetLastError();
}

hipError_t launch_weak(const int *P, const int *nP, int nP_max, const int *cnt, int thr, int *W, int *nW,
                       int *stamp, int tid, hipStream_t st) {
    if (nP_max <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_weak, dim3(blocks_for(nP_max)), dim3(kBlock), 0, st, P, nP, cnt, thr, W, nW, stamp, tid);
    return hipGetLastError();
}

hipError_t launch_path_prune(const DevData &D, uint64_t smask, const int *W, const int *nW, int nW_max,
                             const int *spx, const uint64_t *spm, const int *nsp, int nsp_max, const int *stamp,
                             int tid, unsigned long long *bits, uint64_t *table, uint64_t tmask, int *outx,
                             uint64_t *outm, int *nout, hipStream_t st) {
    if (nsp_max <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill<uint64_t>, dim3(blocks_for(tmask + 1)), dim3(kBlock), 0, st, table, kEmpty,
                       (size_t) (tmask + 1));
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(kBlock), 0, st, bits, 0ull, (size_t) 1);
    hipLaunchKernelGGL(k_path_bits, dim3(blocks_for(nsp_max)), dim3(kBlock), 0, st, spx, spm, nsp, stamp, tid, bits);
    if (nW_max > 0)
        hipLaunchKernelGGL(k_path_insert, dim3(blocks_for(nW_max)), dim3(kBlock), 0, st, D, smask, W, nW, bits, table,
                           tmask);
    hipLaunchKernelGGL(k_path_check, dim3(blocks_for(nsp_max)), dim3(kBlock), 0, st, D, smask, spx, spm, nsp, stamp,
                       tid, table, tmask, outx, outm, nout);
    return hipGetLastError();
}

hipError_t launch_all_queries(const int *cnt, int n, int *queries, int *nq, hipStream_t st) {
    hipLaunchKernelGGL(k_all_queries, dim3(blocks_for(n)), dim3(kBlock), 0, st, cnt, n, queries, nq);
    return hipGetLastError();
}

hipError_t launch_components(const DevData &D, const SubspaceArg &S, float s_star, const int *cnt, int thr,
                             int *parent, int *comp, int *csize, hipStream_t st) {
    const int n = D.n;
    size_t lds = (sizeof(float) + sizeof(int)) * kBlock * S.k;
    hipLaunchKernelGGL(k_uf_init, dim3(blocks_for(n)), dim3(kBlock), 0, st, cnt, n, thr, parent);
    hipLaunchKernelGGL(k_uf_link, dim3(blocks_for(n)), dim3(kBlock), lds, st, D, S, s_star, parent);
    hipLaunchKernelGGL(k_fill<int>, dim3(blocks_for(n)), dim3(kBlock), 0, st, csize, 0, (size_t) n);
    hipLaunchKernelGGL(k_uf_flatten, dim3(blocks_for(n)), dim3(kBlock), 0, st, parent, n, comp, csize);
    return hipGetLastError();
}

hipError_t launch_redundancy(const int *W, const int *nW, int nW_max, int const *const *compH,
                             int const *const *csizeH, uint8_t const *const *inresH, int nH, int *mn,
                             hipStream_t st) {
    if (nW_max <= 0 || nH <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_redundancy, dim3(blocks_for(nW_max), nH), dim3(kBlock), 0, st, W, nW, compH, csizeH, inresH,
                       mn);
    return hipGetLastError();
}

hipError_t launch_seed(const int *W, const int *nW, int nW_max, const int *comp, uint8_t *lab, hipStream_t st) {
    if (nW_max <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_seed, dim3(blocks_for(nW_max)), dim3(kBlock), 0, st, W, nW, comp, lab);
    return hipGetLastError();
}

hipError_t launch_collect(const uint8_t *lab, int n, const int *csize, int *out, int *nout, hipStream_t st) {
    hipLaunchKernelGGL(k_collect, dim3(blocks_for(n)), dim3(kBlock), 0, st, lab, n, csize, out, nout);
    return hipGetLastError();
}

hipError_t launch_join_check(const int *compS, const uint8_t *keep, const int *csizeS, int n, float r,
                             int const *const *compH, int const *const *csizeH, uint8_t const *const *inresH,
                             int nH, uint8_t *drop, hipStream_t st) {
    if (nH <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_join_check, dim3(blocks_for(n), nH), dim3(kBlock), 0, st, compS, keep, csizeS, n, r, compH,
                       csizeH, inresH, drop);
    return hipGetLastError();
}

hipError_t launch_labels(const int *comp, const uint8_t *inres, int n, int *out, hipStream_t st) {
    hipLaunchKernelGGL(k_labels, dim3(blocks_for(n)), dim3(kBlock), 0, st, comp, inres, n, out);
    return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t *p, uint8_t v, size_t count, hipStream_t st) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_fill<uint8_t>, dim3(blocks_for(count)), dim3(kBlock), 0, st, p, v, count);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, int v, size_t count, hipStream_t st) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_fill<int>, dim3(blocks_for(count)), dim3(kBlock), 0, st, p, v, count);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t v, size_t count, hipStream_t st) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_fill<uint64_t>, dim3(blocks_for(count)), dim3(kBlock), 0, st, p, v, count);
    return hipGetLastError();
}

hipError_t launch_set_flags(uint8_t *flags, const int *ids, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_flags, dim3(blocks_for(count)), dim3(kBlock), 0, st, flags, ids, count);
    return hipGetLastError();
}

hipError_t launch_iota(int *p, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(n)), dim3(kBlock), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_root_paths(const uint64_t *nearm, int n, int *outx, uint64_t *outm, int *nout, hipStream_t st) {
    hipLaunchKernelGGL(k_root_paths, dim3(blocks_for(n)), dim3(kBlock), 0, st, nearm, n, outx, outm, nout);
    return hipGetLastError();
}

}  // namespace inscy
