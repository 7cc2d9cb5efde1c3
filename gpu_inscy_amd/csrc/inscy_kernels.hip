// HIP kernels of the MI355X INSCY engine (gfx950, wave64).
//
// Everything here is integer/byte streaming work plus one FP32 distance loop;
// nothing is GEMM-shaped (the contraction depth is |S| <= a few dims), so the
// bound for every kernel is memory bandwidth/latency or VALU issue, not MFMA.
//
// Floating point: the reference compares fl(sqrt(sum_k fl(fl(p_k - q_k)^2))) <= eps
// with separate multiplies and adds in subspace order (ClusteringCpu.cu:40-81).
// This file is compiled with -ffp-contract=off and the pragma below so that no
// multiply-add is fused, and the sqrt is replaced by the exact equivalent
// `sum <= s_star` (s_star = largest float whose sqrtf is <= eps).

#include <climits>
#include <cmath>
#include <cstring>


#include "inscy_device.h"
#include "inscy_internal.h"

#pragma clang fp contract(off)

namespace inscy {

namespace {


// ------------------------------------------------------------ discretisation
// get_cell_no (ScyTreeNode.cu:100-102): min((int)(x / (1/c)), c-1); the
// s-connection test of construct_s_connection (:34-49): x >= (c+1)*(1/c) - eps.
__global__ void k_discretise(const float *__restrict__ X, int n, int d, int cells, float cs, int ecells,
                             float ecs, float eps, float *__restrict__ Xt, uint8_t *__restrict__ cellC,
                             uint16_t *__restrict__ cellE, uint64_t *__restrict__ nearw, int nw,
                             int *__restrict__ bad) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    dmask_t nm = 0;
    bool out_of_range = false;
    for (int k = 0; k < d; ++k) {
        float x = X[(size_t) p * d + k];
        if (!(x >= 0.0f && x <= 1.0f)) out_of_range = true;
        x = fminf(fmaxf(x, 0.0f), 1.0f);  // keeps indices in range; the run is rejected via `bad`
        if (Xt) Xt[(size_t) k * n + p] = x;
        int c = min((int) __fdiv_rn(x, cs), cells - 1);
        int e = min((int) __fdiv_rn(x, ecs), ecells - 1);
        cellC[(size_t) k * n + p] = (uint8_t) c;
        cellE[(size_t) k * n + p] = (uint16_t) e;
        float upper = __fmul_rn((float) (c + 1), cs);
        float bound = __fsub_rn(upper, eps);
        if (x >= bound) nm |= dbit(k);
    }
    // nw 64-bit words per point, dimension k at bit k % 64 of word k / 64
    nearw[(size_t) p * nw] = dlo(nm);
    if (nw > 1) nearw[(size_t) p * nw + 1] = dhi(nm);
    if (out_of_range) atomicOr(bad, 1);
}

// Occupancy of the ceil(1/eps)-grid per dimension; the host derives from it
// which dimensions of a subspace make the most selective grid index.
__global__ void k_hist1(const uint16_t *__restrict__ cellE, int n, int d, int ecells, int *__restrict__ hist) {
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t) n * d) return;
    int k = (int) (i / n);
    atomicAdd(&hist[(size_t) k * ecells + cellE[i]], 1);
}

__global__ void k_histc(const uint8_t *__restrict__ cellC, int n, int d, int cells, int *__restrict__ hist) {
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t) n * d) return;
    int k = (int) (i / n);
    atomicAdd(&hist[(size_t) k * cells + cellC[i]], 1);
}

// ------------------------------------------------------------ expansion of one tree
// mergeWithNeighbors decision input (ScyTreeNode.cu:837-851 with the flag of
// restrict_node :262-267): cell c of dimension b is s-connected iff a path
// branching at b starts from a point in cell c.  One pass sets the flags of
// every dimension b >= first.
__global__ void k_flags_all(const int *__restrict__ spx, const uint64_t *__restrict__ spm, int mw,
                            const int *__restrict__ nsp, const uint8_t *__restrict__ cellC, int n, int first,
                            int cells, int *__restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsp) return;
    int x = spx[i];
    dmask_t m = mload(spm, i, mw) & dabove(first - 1);
    while (m) {
        const int b = dfirst(m);
        m &= m - 1;
        flags[(b - first) * cells + cellC[(size_t) b * n + x]] = 1;
    }
}

// Counting-sort partition of a tree's points, weak-dense points and paths by
// the cell group of dimension m (restrict + mergeWithNeighbors for every
// group at once).  cnt[0..G): points, [G..2G): weak-dense points, [2G..3G):
// paths that branch above m; [3G..6G): scatter cursors.
struct PartArg {
    const int *P;
    const int *nP;
    const int *spx;
    const uint64_t *spm;  // mw words per path
    const int *nsp;
    const uint8_t *cellC_m;
    const int *gid;      // [cells] group of each cell of m
    const int *weakcnt;  // neighbour counts (or phi decision bits) of the subspace S ∪ {m}
    int thr, m, G;
    int nP_blocks, mw, wbit;
};

__global__ void k_part_count(PartArg A, int *__restrict__ cnt) {
    __shared__ int h[3 * 256];
    for (int t = threadIdx.x; t < 3 * A.G; t += blockDim.x) h[t] = 0;
    __syncthreads();
    if ((int) blockIdx.x < A.nP_blocks) {
        int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < *A.nP) {
            int p = A.P[i];
            int g = A.gid[A.cellC_m[p]];
            atomicAdd(&h[g], 1);
            if (meets(A.weakcnt[p], A.thr, A.wbit)) atomicAdd(&h[A.G + g], 1);
        }
    } else {
        int i = (blockIdx.x - A.nP_blocks) * blockDim.x + threadIdx.x;
        if (i < *A.nsp) {
            const dmask_t mk = mload(A.spm, i, A.mw) & dabove(A.m);
            if (mk) atomicAdd(&h[2 * A.G + A.gid[A.cellC_m[A.spx[i]]]], 1);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * A.G; t += blockDim.x)
        if (h[t]) atomicAdd(&cnt[t], h[t]);
}

__global__ void k_part_scatter(PartArg A, int *__restrict__ cnt, int *__restrict__ outP, int *__restrict__ outW,
                               int *__restrict__ outx, uint64_t *__restrict__ outm) {
    __shared__ int base[3 * 256];
    if (threadIdx.x < 3) {
        int acc = 0;
        for (int g = 0; g < A.G; ++g) {
            base[threadIdx.x * A.G + g] = acc;
            acc += cnt[threadIdx.x * A.G + g];
        }
    }
    __syncthreads();
    int *cur = cnt + 3 * A.G;
    if ((int) blockIdx.x < A.nP_blocks) {
        int i = blockIdx.x * blockDim.x + threadIdx.x;
        bool in = i < *A.nP;
        int p = in ? A.P[i] : 0;
        int g = in ? A.gid[A.cellC_m[p]] : 0;
        bool weak = in && meets(A.weakcnt[p], A.thr, A.wbit);
        int sp = wave_add_by_key(cur, g, in);
        int sw = wave_add_by_key(cur + A.G, g, weak);
        if (in) outP[base[g] + sp] = p;
        if (weak) outW[base[A.G + g] + sw] = p;
    } else {
        int i = (blockIdx.x - A.nP_blocks) * blockDim.x + threadIdx.x;
        dmask_t mk = 0;
        int x = 0, g = 0;
        if (i < *A.nsp) {
            mk = mload(A.spm, i, A.mw) & dabove(A.m);
            x = A.spx[i];
            g = A.gid[A.cellC_m[x]];
        }
        int sx = wave_add_by_key(cur + 2 * A.G, g, mk != 0);
        if (mk) {
            int pos = base[2 * A.G + g] + sx;
            outx[pos] = x;
            mstore(outm, pos, mk, A.mw);
        }
    }
}

// ------------------------------------------------------------ epsilon neighbourhood: subspace grid pass
// Candidate rule of get_possible_neighbors_from (ScyTreeNode.cu:896-946) on
// the ceil(1/eps)-grid: |cell_e(q) - cell_e(p)| <= 1 in every dimension of S;
// then the float distance in S (ClusteringCpu.cu:52-81).  Per subspace the
// points are counting-sorted by their cells in the j most selective
// dimensions of S (j <= 3), so the candidates of a query are 3^(j-1)
// contiguous ranges of the sorted order; the coordinates and cells of S are
// gathered into that order ([k][n]).  Queries run in the same sorted order and
// kLanesPerQuery lanes stride through each query's ranges, so the lanes of a
// wave read consecutive candidates and neighbouring queries share ranges.


__global__ void k_grid_keys(DevData D, SubspaceArg S, GridArg G, int *__restrict__ keys, int *__restrict__ rank,
                            int *__restrict__ hist) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.n) return;
    int key = 0;
    for (int a = 0; a < G.j; ++a) key = key * G.C + D.cellE[(size_t) S.dims[G.tI[a]] * D.n + p];
    keys[p] = key;
    rank[p] = atomicAdd(&hist[key], 1);
}

// Sorted order and the subspace's coordinates/cells gathered into it.
__global__ void k_grid_scatter(DevData D, SubspaceArg S, const int *__restrict__ keys, const int *__restrict__ rank,
                               const int *__restrict__ off, int *__restrict__ order, float *__restrict__ Ys,
                               uint16_t *__restrict__ Es) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.n) return;
    const int n = D.n;
    int pos = off[keys[p]] + rank[p];
    order[pos] = p;
    for (int t = 0; t < S.k; ++t) {
        int s = S.dims[t];
        Ys[(size_t) t * n + pos] = D.Xt[(size_t) s * n + p];
        Es[(size_t) t * n + pos] = D.cellE[(size_t) s * n + p];
    }
}


// |N_S(p)| of the points still uncounted (cnt == -1) and, when `mark` is
// given, stamped with `markid`.  Counting stops once `cap` neighbours are
// found: every comparison the algorithm makes with a count is against a
// threshold <= cap, so min(|N_S(p)|, cap)-or-more decides them all.
__global__ __launch_bounds__(kBlock) void k_grid_count(int n, SubspaceArg S, GridArg G, const int *__restrict__ off,
                                                       const int *__restrict__ order, const float *__restrict__ Ys,
                                                       const uint16_t *__restrict__ Es, float s_star, int cap,
                                                       const int *__restrict__ mark, int markid,
                                                       const int *__restrict__ bound, int lo, int *__restrict__ cnt,
                                                       unsigned long long *__restrict__ stat) {
    constexpr int LPQ = kLanesPerQuery, QPB = kBlock / LPQ;
    extern __shared__ char smem[];
    float *sy = reinterpret_cast<float *>(smem);
    int *se = reinterpret_cast<int *>(smem + sizeof(float) * QPB * S.k);
    const int lq = threadIdx.x / LPQ, sub = threadIdx.x % LPQ;
    const int i = blockIdx.x * QPB + lq;
    int p = i < n ? order[i] : 0;
    bool active = i < n && cnt[p] < 0 && (!mark || mark[p] == markid);
    // N_S(p) is a subset of N_S'(p) for S' a subset of S (the rounded distance
    // sum only grows with more dimensions), so a complete count of a subspace
    // of S below both thresholds of S decides p without a scan.
    if (active && bound) {
        int b = bound[p];
        if (b >= 0 && b < lo) {
            if (sub == 0) cnt[p] = b;
            active = false;
        }
    }
    if (sub == 0 && active) stage_query(S, n, Ys, Es, i, sy + lq, se + lq, QPB);
    __syncthreads();
    long long seen = 0;
    int found = 0;
    if (active) {
        const float *qy = sy + lq;
        const int *qe = se + lq;
        const int gshift = (__lane_id() / LPQ) * LPQ;
        for_ranges(G, off, qe, QPB, [&](int a, int b) {
            for (int base = a; base < b && found < cap; base += LPQ) {
                int jj = base + sub;
                bool hit = jj < b && jj != i && is_neighbour(S, G, n, Ys, Es, qy, qe, QPB, jj, s_star);
                found += __popcll((__ballot(hit) >> gshift) & ((1ull << LPQ) - 1ull));
                seen += min(LPQ, b - base);
            }
        });
        if (sub == 0) cnt[p] = found;
    }
    if (stat) {
        // per wave, into one of 64 slot groups (16 words apart): [0] candidates
        // examined, [1] points counted, [2] algorithmic bytes = per point the
        // query's k coordinates and its count (4k + 4) plus, per neighbour
        // established, its id and k coordinates (4 + 4k)
        const bool mine = sub == 0 && active;
        long long done = mine ? 1 : 0;
        long long abytes = mine ? (4ll * S.k + 4) + (long long) found * (4 + 4ll * S.k) : 0;
        if (sub != 0) seen = 0;
        for (int o = 32; o > 0; o >>= 1) {
            seen += __shfl_down(seen, o);
            done += __shfl_down(done, o);
            abytes += __shfl_down(abytes, o);
        }
        if (__lane_id() == 0 && done) {
            unsigned long long *slot = stat + 16 * (blockIdx.x & 63);
            atomicAdd(&slot[0], (unsigned long long) seen);
            atomicAdd(&slot[1], (unsigned long long) done);
            atomicAdd(&slot[2], (unsigned long long) abytes);
        }
    }
}

// phi mode (rectangular = false): the decision bits of the points without a
// value (cnt == -1; stamped markid when mark is given).  The lanes of a query
// sum its phi terms in double in scan order; a threshold outside the rounding
// band of the reference's float leaf-order sum (phirel per term) is decided
// here, and a point with a threshold inside it goes to fb for the leaf-order
// recomputation.  Stored like the level-synchronous passes: phi_code(bits, U).
__global__ __launch_bounds__(kBlock) void k_grid_phi(int n, SubspaceArg S, GridArg G, const int *__restrict__ off,
                                                     const int *__restrict__ order, const float *__restrict__ Ys,
                                                     const uint16_t *__restrict__ Es, float s_star, float eps,
                                                     double phirel, float phi_w, float phi_d,
                                                     const int *__restrict__ mark, int markid, int *__restrict__ cnt,
                                                     int2 *__restrict__ fb, int *__restrict__ nfb, int fbcap) {
    constexpr int LPQ = kLanesPerQuery, QPB = kBlock / LPQ;
    extern __shared__ char smem[];
    float *sy = reinterpret_cast<float *>(smem);
    int *se = reinterpret_cast<int *>(smem + sizeof(float) * QPB * S.k);
    const int lq = threadIdx.x / LPQ, sub = threadIdx.x % LPQ;
    const int i = blockIdx.x * QPB + lq;
    const int p = i < n ? order[i] : 0;
    const bool active = i < n && cnt[p] < 0 && (!mark || mark[p] == markid);
    if (sub == 0 && active) stage_query(S, n, Ys, Es, i, sy + lq, se + lq, QPB);
    __syncthreads();
    if (!active) return;  // the walk's ballots below span only active query groups
    const float *qy = sy + lq;
    const int *qe = se + lq;
    const int gshift = (__lane_id() / LPQ) * LPQ;
    int found = 0;
    double dsum = 0.0;
    for_ranges(G, off, qe, QPB, [&](int a, int b) {
        for (int base = a; base < b; base += LPQ) {
            const int jj = base + sub;
            float acc = 0.0f;
            bool hit = jj < b && jj != i;
            for (int t = 0; t < S.k && hit; ++t) {
                if ((G.checkmask >> t) & 1ull) {
                    const int de = (int) Es[(size_t) t * n + jj] - qe[t * QPB];
                    if (de < -1 || de > 1) hit = false;
                }
                if (hit) {
                    const float diff = qy[t * QPB] - Ys[(size_t) t * n + jj];
                    const float sq = diff * diff;
                    acc = acc + sq;
                    if (acc > s_star) hit = false;
                }
            }
            if (hit) dsum += phi_term(acc, eps);
            found += __popcll((__ballot(hit) >> gshift) & ((1ull << LPQ) - 1ull));
        }
    });
    for (int o = LPQ / 2; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o);
    if (sub != 0) return;
    const double lo = dsum * (1.0 - (double) found * phirel);
    const double hi = dsum * (1.0 + (double) found * phirel);
    const double tw = (double) phi_w, td = (double) phi_d;
    int code = 0;
    bool unsure = false;
    if (lo >= tw) code |= 1;
    else if (!(hi < tw)) unsure = true;
    if (lo >= td) code |= 2;
    else if (!(hi < td)) unsure = true;
    cnt[p] = phi_code(code, hi * (1.0 + (double) found * phirel));
    if (unsure) {
        const int at = atomicAdd(nfb, 1);
        if (at < fbcap) fb[at] = make_int2(0, p);
    }
}

// dense flag per sorted position, union-find roots, zeroed component sizes
__global__ void k_dense_init(const int *__restrict__ order, int n, const int *__restrict__ cnt, int thr_dense,
                             int dbit, uint8_t *__restrict__ dsorted, int *__restrict__ par, int *__restrict__ csize) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // i: a sorted position for the dense flag, a point id for the rest
    dsorted[i] = meets(cnt[order[i]], thr_dense, dbit) ? 1 : 0;
    par[i] = meets(cnt[i], thr_dense, dbit) ? i : -1;
    csize[i] = 0;
}

// ------------------------------------------------------------ dense components
// Union-find over the dense-dense epsilon graph; hooking the larger root under
// the smaller keeps every root the smallest point id of its component, which
// is also the canonical cluster name.

// Components in the Afforest manner (Sutton, Ben-Nun, Bar-Noy, "Optimizing
// Parallel Graph Connectivity Computation via Subgraph Sampling", IPDPS 2018):
// (1) every dense point links to its first two dense neighbours, (2) roots are
// compressed, (3) the component holding most of a fixed sample of points is
// taken as the giant one, (4) only points outside it link to all their dense
// neighbours (both directions, so an edge between the giant component and the
// rest is still seen from the outside end).
__global__ __launch_bounds__(kBlock) void k_link_sample(int n, SubspaceArg S, GridArg G, const int *__restrict__ off,
                                                        const int *__restrict__ order, const float *__restrict__ Ys,
                                                        const uint16_t *__restrict__ Es,
                                                        const uint8_t *__restrict__ dsorted, float s_star,
                                                        int *__restrict__ par) {
    constexpr int LPQ = kLanesPerQuery, QPB = kBlock / LPQ;
    extern __shared__ char smem[];
    float *sy = reinterpret_cast<float *>(smem);
    int *se = reinterpret_cast<int *>(smem + sizeof(float) * QPB * S.k);
    const int lq = threadIdx.x / LPQ, sub = threadIdx.x % LPQ;
    const int i = blockIdx.x * QPB + lq;
    const bool active = i < n && dsorted[i];
    if (sub == 0 && active) stage_query(S, n, Ys, Es, i, sy + lq, se + lq, QPB);
    __syncthreads();
    if (!active) return;
    const float *qy = sy + lq;
    const int *qe = se + lq;
    const int p = order[i];
    const int gshift = (__lane_id() / LPQ) * LPQ;  // this query's lanes within the wave
    const unsigned long long below = (1ull << sub) - 1ull;
    int found = 0;
    for_ranges(G, off, qe, QPB, [&](int a, int b) {
        for (int base = a; base < b && found < 2; base += LPQ) {
            int jj = base + sub;
            bool hit = jj < b && jj != i && dsorted[jj] && is_neighbour(S, G, n, Ys, Es, qy, qe, QPB, jj, s_star);
            unsigned long long mine = (__ballot(hit) >> gshift) & ((1ull << LPQ) - 1ull);
            if (hit && __popcll(mine & below) < 2 - found) uf_union(par, p, order[jj]);
            found += min(__popcll(mine), 2 - found);
        }
    });
}

__global__ void k_uf_compress(int *__restrict__ par, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && par[i] >= 0) par[i] = uf_find(par, i);
}

// one block of 256: the most frequent root among 256 evenly spaced points
__global__ void k_giant(const int *__restrict__ par, int n, int *__restrict__ giant) {
    __shared__ int roots[256];
    __shared__ int best_c[256];
    __shared__ int best_r[256];
    const int t = threadIdx.x;
    int p = (int) (((long long) t * n) / 256);
    roots[t] = p < n ? par[p] : -1;
    __syncthreads();
    int r = roots[t], c = 0;
    if (r >= 0)
        for (int u = 0; u < 256; ++u) c += roots[u] == r;
    best_c[t] = c;
    best_r[t] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w && (best_c[t + w] > best_c[t] || (best_c[t + w] == best_c[t] && best_r[t + w] < best_r[t] &&
                                                    best_c[t + w] > 0))) {
            best_c[t] = best_c[t + w];
            best_r[t] = best_r[t + w];
        }
        __syncthreads();
    }
    if (t == 0) *giant = best_c[0] > 0 ? best_r[0] : -1;
}

__global__ __launch_bounds__(kBlock) void k_link_full(int n, SubspaceArg S, GridArg G, const int *__restrict__ off,
                                                      const int *__restrict__ order, const float *__restrict__ Ys,
                                                      const uint16_t *__restrict__ Es,
                                                      const uint8_t *__restrict__ dsorted, float s_star,
                                                      const int *__restrict__ giant, int *__restrict__ par) {
    constexpr int LPQ = kLanesPerQuery, QPB = kBlock / LPQ;
    extern __shared__ char smem[];
    float *sy = reinterpret_cast<float *>(smem);
    int *se = reinterpret_cast<int *>(smem + sizeof(float) * QPB * S.k);
    const int lq = threadIdx.x / LPQ, sub = threadIdx.x % LPQ;
    const int i = blockIdx.x * QPB + lq;
    bool active = i < n && dsorted[i];
    int p = active ? order[i] : 0;
    if (active && par[p] == *giant) active = false;
    if (sub == 0 && active) stage_query(S, n, Ys, Es, i, sy + lq, se + lq, QPB);
    __syncthreads();
    if (!active) return;
    const float *qy = sy + lq;
    const int *qe = se + lq;
    int rp = uf_find(par, p);
    for_ranges(G, off, qe, QPB, [&](int a, int b) {
        for (int jj = a + sub; jj < b; jj += LPQ) {
            if (jj == i || !dsorted[jj] || !is_neighbour(S, G, n, Ys, Es, qy, qe, QPB, jj, s_star)) continue;
            int q = order[jj];
            if (uf_find(par, q) == rp) continue;
            uf_union(par, p, q);
            rp = uf_find(par, p);
        }
    });
}

__global__ void k_uf_flatten(int *__restrict__ par, int n, int *__restrict__ comp, int *__restrict__ csize) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1;
    if (i < n && par[i] >= 0) c = uf_find(par, i);
    wave_add_by_key(csize, c, c >= 0);
    if (i < n) comp[i] = c;
}

// ------------------------------------------------------------ s-connection paths after the prune
// A path (b, x) survives the prune of tree R iff some surviving point w shares
// x's cells on every unrestricted dimension <= b.  (propergate_count2,
// ScyTreeNode.cu:446-467, erases the s-connection whose same-cell regular
// sibling lost all its points; the regular node a path hangs from keeps a
// point whenever that sibling does.)  Membership is answered by an
// open-addressing table of 64-bit entries (32 hash bits | representative point
// id); every hash match is confirmed by comparing the cells themselves, so the
// test is exact.


__global__ void k_stamp(const int *__restrict__ W, const int *__restrict__ nW, int *__restrict__ stamp, int tid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *nW) stamp[W[i]] = tid;
}

__global__ void k_path_bits(const int *__restrict__ spx, const uint64_t *__restrict__ spm, int mw,
                            const int *__restrict__ nsp, const int *__restrict__ stamp, int tid,
                            dmask_t *__restrict__ bits) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    dmask_t m = 0;
    if (i < *nsp && stamp[spx[i]] != tid) m = mload(spm, i, mw);
    wave_or_mask(m, bits);
}

__global__ void k_path_insert(DevData D, dmask_t smask, const int *__restrict__ W, const int *__restrict__ nW,
                              const dmask_t *__restrict__ bits, uint64_t *__restrict__ table, uint64_t tmask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nW) return;
    const dmask_t need = *bits;
    if (!need) return;
    int w = W[i];
    uint64_t h = 0x243F6A8885A308D3ull;
    for (int t = 0; t < D.d; ++t) {
        const int k = D.level[t];
        if (dhas(smask, k)) continue;
        h = mix64(h + D.cellC[(size_t) k * D.n + w] + 1);
        if (!dhas(need, k)) continue;
        uint64_t e = mix64(h ^ (0x9E3779B97F4A7C15ull * (uint64_t) (k + 1)));
        uint64_t entry = (e & 0xFFFFFFFF00000000ull) | (uint32_t) w;
        uint64_t slot = e & tmask;
        while (true) {
            // a slot is written once, by the swap from empty: a plain load that
            // still sees it empty is corrected by the failed swap, a full one is final
            uint64_t cur = table[slot];
            if (cur == kEmpty) {
                cur = atomicCAS((unsigned long long *) &table[slot], (unsigned long long) kEmpty,
                                (unsigned long long) entry);
                if (cur == kEmpty) break;
            }
            if ((cur >> 32) == (entry >> 32) && same_prefix(D, smask, w, (int) (uint32_t) cur, t)) break;
            slot = (slot + 1) & tmask;
        }
    }
}

__global__ void k_path_check(DevData D, dmask_t smask, const int *__restrict__ spx, const uint64_t *__restrict__ spm,
                             int mw,
                             const int *__restrict__ nsp, const int *__restrict__ stamp, int tid,
                             const uint64_t *__restrict__ table, uint64_t tmask, int *__restrict__ outx,
                             uint64_t *__restrict__ outm, int *__restrict__ nout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int x = 0;
    dmask_t m = 0;
    if (i < *nsp) {
        x = spx[i];
        m = mload(spm, i, mw);
        if (stamp[x] != tid) {
            uint64_t h = 0x243F6A8885A308D3ull;
            dmask_t keep = 0;
            for (int t = 0; t < D.d; ++t) {
                const int k = D.level[t];
                if (dhas(smask, k)) continue;
                h = mix64(h + D.cellC[(size_t) k * D.n + x] + 1);
                if (!dhas(m, k)) continue;
                uint64_t e = mix64(h ^ (0x9E3779B97F4A7C15ull * (uint64_t) (k + 1)));
                uint64_t slot = e & tmask;
                while (true) {
                    uint64_t cur = table[slot];  // complete: the inserts ran in an earlier launch
                    if (cur == kEmpty) break;
                    if ((cur >> 32) == (e >> 32) && same_prefix(D, smask, x, (int) (uint32_t) cur, t)) {
                        keep |= dbit(k);
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
        mstore(outm, slot, m, mw);
    }
}

// ------------------------------------------------------------ redundancy and join on the device
// The result lives on the device as a table of subspace entries; an entry is
// in the result while nkept > 0.  The decisions run in stream order, which is
// the reference's depth-first order, without any host round trip.

// entries e whose mask contains smask (subspace_of, util.cu:619-631) and that
// are in the result now
__global__ void k_select(const Entry *__restrict__ E, const int *__restrict__ nE, dmask_t smask, int *__restrict__ list,
                         int *__restrict__ nlist) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int ne = *nE;
    for (int e = threadIdx.x; e < ne; e += blockDim.x)
        if ((E[e].mask & smask) == smask && E[e].nkept > 0) list[atomicAdd(&cnt, 1)] = e;
    __syncthreads();
    if (threadIdx.x == 0) *nlist = cnt;
}

// pruneRedundancy (ScyTreeNode.cu:669-734): per listed superspace H, the
// smallest result cluster of H holding a point of the region.  The members of
// H's result clusters are listed, the region's points are stamped with `tid`.
__global__ void k_redundancy(const int *__restrict__ stamp, int tid, const Entry *__restrict__ E,
                             const int *__restrict__ list, const int *__restrict__ nlist, int *__restrict__ mn) {
    const int nl = *nlist;
    const int stride = gridDim.x * blockDim.x;
    for (int h = 0; h < nl; ++h) {
        const Entry e = E[list[h]];
        int best = INT_MAX;
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < e.nmem; t += stride) {
            int q = e.mem[t];
            if (stamp[q] == tid) best = min(best, e.csize[e.comp[q]]);
        }
        for (int o = 32; o > 0; o >>= 1) best = min(best, __shfl_down(best, o));
        if (__lane_id() == 0 && best != INT_MAX) atomicMin(&mn[h], best);
    }
}

// keep iff no superspace cluster touches the region or |W| * r > max_h min_h
__global__ void k_decide(const int *__restrict__ nlist, int *__restrict__ mn, int nW, float r,
                         int *__restrict__ keep) {
    if (threadIdx.x != 0) return;
    const int nl = *nlist;
    int max_min = 0;
    for (int h = 0; h < nl; ++h) {
        if (mn[h] != INT_MAX && mn[h] > max_min) max_min = mn[h];
        mn[h] = INT_MAX;
    }
    float lhs = (float) nW * r;
    *keep = (max_min == 0 || (double) lhs > (double) max_min) ? 1 : 0;
}

// cluster (ClusteringCpu.cu:317-389): a BFS from region point w labels w's
// dense component; here the component is marked labelled.
__global__ void k_seed_if(const int *__restrict__ W, int nW, const int *__restrict__ comp, const int *__restrict__ keep,
                          uint8_t *__restrict__ lab) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nW || !*keep) return;
    int c = comp[W[i]];
    if (c >= 0) lab[c] = 1;
}

// join (util.cu:655-743), step 1: labelled components of at least min_size
__global__ void k_join_keep(const uint8_t *__restrict__ lab, int n, const int *__restrict__ csize, int min_size,
                            uint8_t *__restrict__ keepf, uint8_t *__restrict__ drop) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    keepf[k] = (lab[k] && csize[k] >= min_size) ? 1 : 0;
    drop[k] = 0;
}

// step 2: drop K when one of its points lies in a result cluster of a
// superspace H larger than |K| * r (walking the members of H's clusters)
__global__ void k_join_check(const int *__restrict__ compS, const uint8_t *__restrict__ keepf,
                             const int *__restrict__ csizeS, float r, const Entry *__restrict__ E,
                             const int *__restrict__ list, const int *__restrict__ nlist, uint8_t *__restrict__ drop) {
    const int nl = *nlist;
    const int stride = gridDim.x * blockDim.x;
    for (int h = 0; h < nl; ++h) {
        const Entry e = E[list[h]];
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < e.nmem; t += stride) {
            int q = e.mem[t];
            int K = compS[q];
            if (K < 0 || !keepf[K]) continue;
            if ((float) csizeS[K] * r < (float) e.csize[e.comp[q]]) drop[K] = 1;
        }
    }
}

// step 3: insert the survivors into the result of S
__global__ void k_join_apply(const uint8_t *__restrict__ keepf, const uint8_t *__restrict__ drop, int n,
                             uint8_t *__restrict__ inres, uint8_t *__restrict__ fresh, Entry *__restrict__ E, int eS) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int added = 0;
    if (k < n) {
        fresh[k] = 0;
        if (keepf[k] && !drop[k] && !inres[k]) {
            inres[k] = 1;
            fresh[k] = 1;
            added = 1;
        }
    }
    for (int o = 32; o > 0; o >>= 1) added += __shfl_down(added, o);
    if (__lane_id() == 0 && added) atomicAdd(&E[eS].nkept, added);
}

// step 4: append the members of the newly inserted clusters to S's member list
__global__ void k_join_members(const int *__restrict__ compS, int n, const uint8_t *__restrict__ fresh,
                               Entry *__restrict__ E, int eS) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (i < n) {
        int K = compS[i];
        in = K >= 0 && fresh[K];
    }
    int slot = wave_slot(in, &E[eS].nmem);
    if (in) E[eS].mem[slot] = i;
}

__global__ void k_set_entry(Entry *__restrict__ E, int idx, Entry e, int *__restrict__ nE, int count) {
    if (threadIdx.x == 0) {
        E[idx] = e;
        *nE = count;
    }
}

// Canonical labels of a chunk of result subspaces (grid.y = row): clusters
// are numbered by their smallest point id, which is the component root.
// Pass 1 flags the kept roots, an exclusive scan over the whole chunk numbers
// them, pass 2 writes every point's number relative to its row's first root.
__global__ void k_emit_roots(const EmitRow *__restrict__ rows, int n, int *__restrict__ flags) {
    const EmitRow R = rows[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[(size_t) blockIdx.y * n + i] = (R.comp[i] == i && R.inres[i]) ? 1 : 0;
}

// T = int: the label; T = uint8_t / uint16_t: the label + 1 (0 for noise), for
// rows whose labels fit (the host reads them back at a quarter or half the bytes)
template <class T>
__global__ void k_emit_labels(const EmitRow *__restrict__ rows, int n, const int *__restrict__ idx,
                              T *__restrict__ labels) {
    const EmitRow R = rows[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t row = (size_t) blockIdx.y * n;
    const int c = R.comp[i];
    const int v = (c >= 0 && R.inres[c]) ? idx[row + c] - idx[row] : -1;
    labels[row + i] = (T) (sizeof(T) == sizeof(int) ? v : v + 1);
}

__global__ void k_emit_counts(const int *__restrict__ idx, int n, int r, int *__restrict__ counts) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < r) counts[t] = idx[(size_t) (t + 1) * n] - idx[(size_t) t * n];
}

template <class T>
__global__ void k_fill(T *p, T v, size_t count) {
    fill_span(p, v, count);
}

__global__ void k_gather(const int *__restrict__ src, const int *__restrict__ idx, int count, int *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = src[idx[i]];
}

__global__ void k_iota(int *__restrict__ p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_rank_keys(const uint16_t *__restrict__ cellE, int n, int dim, const int *__restrict__ perm,
                            uint16_t *__restrict__ keys) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = cellE[(size_t) dim * n + perm[i]];
}

// Root s-connection paths: every point with at least one near-edge dimension.
__global__ void k_root_paths(const dmask_t *__restrict__ nearm, int n, int *__restrict__ outx,
                             uint64_t *__restrict__ outm, int mw, int *__restrict__ nout) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const dmask_t m = i < n ? nearm[i] : (dmask_t) 0;
    int slot = wave_slot(m != 0, nout);
    if (m != 0) {
        outx[slot] = i;
        mstore(outm, slot, m, mw);
    }
}

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

#define LAUNCH(kernel, grid, block, lds, st, ...) hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__)

hipError_t launch_discretise(const float *X, int n, int d, int cells, float eps, float *Xt, uint8_t *cellC,
                             uint16_t *cellE, uint64_t *nearw, int nw, int *bad, hipStream_t st) {
    int ecells = (int) ceil(1.0 / (double) eps);
    float cs = 1.0f / (float) cells;
    float ecs = 1.0f / (float) ecells;
    LAUNCH(k_discretise, dim3(blocks_for(n)), dim3(kBlock), 0, st, X, n, d, cells, cs, ecells, ecs, eps, Xt, cellC,
           cellE, nearw, nw, bad);
    return hipGetLastError();
}

hipError_t launch_hist1(const uint16_t *cellE, int n, int d, int ecells, int *hist, hipStream_t st) {
    size_t nd = (size_t) n * d;
    LAUNCH(k_fill<int>, dim3(fill_blocks((size_t) d * ecells, sizeof(int))), dim3(kBlock), 0, st, hist, 0, (size_t) d * ecells);
    LAUNCH(k_hist1, dim3(blocks_for(nd)), dim3(kBlock), 0, st, cellE, n, d, ecells, hist);
    return hipGetLastError();
}

hipError_t launch_histc(const uint8_t *cellC, int n, int d, int cells, int *hist, hipStream_t st) {
    size_t nd = (size_t) n * d;
    LAUNCH(k_fill<int>, dim3(fill_blocks((size_t) d * cells, sizeof(int))), dim3(kBlock), 0, st, hist, 0, (size_t) d * cells);
    LAUNCH(k_histc, dim3(blocks_for(nd)), dim3(kBlock), 0, st, cellC, n, d, cells, hist);
    return hipGetLastError();
}


static hipError_t build_grid(const DevData &D, const SubspaceArg &S, const GridArg &G, const GridBuffers &B,
                             hipStream_t st) {
    const int n = D.n;
    int ncell = 1;
    for (int a = 0; a < G.j; ++a) ncell *= G.C;
    LAUNCH(k_fill<int>, dim3(fill_blocks((size_t) ncell + 1, sizeof(int))), dim3(kBlock), 0, st, B.hist, 0, (size_t) ncell + 1);
    LAUNCH(k_grid_keys, dim3(blocks_for(n)), dim3(kBlock), 0, st, D, S, G, B.keys, B.rank, B.hist);
    hipError_t e = launch_exclusive_sum(B.scan_temp, B.scan_bytes, B.hist, B.off, ncell + 1, st);
    if (e != hipSuccess) return e;
    LAUNCH(k_grid_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, st, D, S, B.keys, B.rank, B.off, B.order, B.Ys, B.Es);
    return hipGetLastError();
}

static unsigned query_blocks(int n) {
    return (unsigned) ((n + kBlock / kLanesPerQuery - 1) / (kBlock / kLanesPerQuery));
}

static size_t query_lds(const SubspaceArg &S) { return (sizeof(float) + sizeof(int)) * (kBlock / kLanesPerQuery) * S.k; }

hipError_t launch_subspace_counts(const DevData &D, const SubspaceArg &S, const GridArg &G, float s_star, int cap,
                                  const GridBuffers &B, const int *mark, int markid, int *cnt,
                                  unsigned long long *stat, const KernelTimer &tm, hipStream_t st) {
    hipError_t e = build_grid(D, S, G, B, st);
    if (e != hipSuccess) return e;
    if (tm.start) (void) hipEventRecord(tm.start, st);
    LAUNCH(k_grid_count, dim3(query_blocks(D.n)), dim3(kBlock), query_lds(S), st, D.n, S, G, B.off, B.order, B.Ys,
           B.Es, s_star, cap, mark, markid, (const int *) nullptr, 0, cnt, stat);
    if (tm.stop) (void) hipEventRecord(tm.stop, st);
    return hipGetLastError();
}

hipError_t launch_subspace_phi(const DevData &D, const SubspaceArg &S, const GridArg &G, const GridBuffers &B,
                              float s_star, float eps, double phirel, float phi_w, float phi_d, const int *mark,
                              int markid, int *cnt, int2 *fb, int *nfb, int fbcap, hipStream_t st) {
    hipError_t e = build_grid(D, S, G, B, st);
    if (e != hipSuccess) return e;
    LAUNCH(k_fill<int>, dim3(fill_blocks((size_t) 1, sizeof(int))), dim3(kBlock), 0, st, nfb, 0, (size_t) 1);
    LAUNCH(k_grid_phi, dim3(query_blocks(D.n)), dim3(kBlock), query_lds(S), st, D.n, S, G, B.off, B.order, B.Ys, B.Es,
           s_star, eps, phirel, phi_w, phi_d, mark, markid, cnt, fb, nfb, fbcap);
    return hipGetLastError();
}

hipError_t launch_subspace_components(const DevData &D, const SubspaceArg &S, const GridArg &G, float s_star,
                                      int cap, int lo, int thr_dense, int dbit, const GridBuffers &B, const int *bound,
                                      int *cnt, int *comp, int *csize, unsigned long long *stat,
                                      const KernelTimer &tm, hipStream_t st) {
    hipError_t e = build_grid(D, S, G, B, st);
    if (e != hipSuccess) return e;
    const int n = D.n;
    if (tm.start) (void) hipEventRecord(tm.start, st);
    LAUNCH(k_grid_count, dim3(query_blocks(n)), dim3(kBlock), query_lds(S), st, n, S, G, B.off, B.order, B.Ys, B.Es,
           s_star, cap, (const int *) nullptr, 0, bound, lo, cnt, stat);
    if (tm.stop) (void) hipEventRecord(tm.stop, st);
    LAUNCH(k_dense_init, dim3(blocks_for(n)), dim3(kBlock), 0, st, B.order, n, cnt, thr_dense, dbit, B.dsorted, B.par,
           csize);
    LAUNCH(k_link_sample, dim3(query_blocks(n)), dim3(kBlock), query_lds(S), st, n, S, G, B.off, B.order, B.Ys, B.Es,
           B.dsorted, s_star, B.par);
    LAUNCH(k_uf_compress, dim3(blocks_for(n)), dim3(kBlock), 0, st, B.par, n);
    LAUNCH(k_giant, dim3(1), dim3(256), 0, st, B.par, n, B.giant);
    LAUNCH(k_link_full, dim3(query_blocks(n)), dim3(kBlock), query_lds(S), st, n, S, G, B.off, B.order, B.Ys, B.Es,
           B.dsorted, s_star, B.giant, B.par);
    LAUNCH(k_uf_flatten, dim3(blocks_for(n)), dim3(kBlock), 0, st, B.par, n, comp, csize);
    return hipGetLastError();
}

hipError_t launch_mark(const int *P, const int *nP, int nP_max, int *mark, int markid, hipStream_t st) {
    if (nP_max <= 0) return hipSuccess;
    LAUNCH(k_stamp, dim3(blocks_for(nP_max)), dim3(kBlock), 0, st, P, nP, mark, markid);
    return hipGetLastError();
}

hipError_t launch_flags_all(const int *spx, const uint64_t *spm, int mw, const int *nsp, int nsp_max,
                            const uint8_t *cellC, int n, int first, int cells, int *flags, hipStream_t st) {
    if (nsp_max <= 0) return hipSuccess;
    LAUNCH(k_flags_all, dim3(blocks_for(nsp_max)), dim3(kBlock), 0, st, spx, spm, mw, nsp, cellC, n, first, cells,
           flags);
    return hipGetLastError();
}

hipError_t launch_partition(const int *P, const int *nP, int nP_max, const int *spx, const uint64_t *spm, int mw,
                            const int *nsp, int nsp_max, const uint8_t *cellC_m, const int *gid, int G,
                            const int *weakcnt, int thr, int wbit, int m, int *cnt, int *outP, int *outW, int *outx,
                            uint64_t *outm, hipStream_t st) {
    int bp = (int) blocks_for(nP_max), bs = (int) blocks_for(nsp_max);
    LAUNCH(k_fill<int>, dim3(fill_blocks((size_t) 6 * G, sizeof(int))), dim3(kBlock), 0, st, cnt, 0, (size_t) 6 * G);
    if (bp + bs == 0) return hipGetLastError();
    PartArg A{P, nP, spx, spm, nsp, cellC_m, gid, weakcnt, thr, m, G, bp, mw, wbit};
    LAUNCH(k_part_count, dim3(bp + bs), dim3(kBlock), 0, st, A, cnt);
    LAUNCH(k_part_scatter, dim3(bp + bs), dim3(kBlock), 0, st, A, cnt, outP, outW, outx, outm);
    return hipGetLastError();
}

hipError_t launch_path_prune(const DevData &D, dmask_t smask, const int *W, const int *nW, int nW_max,
                             const int *spx, const uint64_t *spm, int mw, const int *nsp, int nsp_max, int *stamp,
                             int tid, dmask_t *bits, uint64_t *table, uint64_t tmask, int *outx, uint64_t *outm,
                             int *nout, hipStream_t st) {
    LAUNCH(k_fill<int>, dim3(fill_blocks((size_t) 1, sizeof(int))), dim3(kBlock), 0, st, nout, 0, (size_t) 1);
    if (nsp_max <= 0) return hipGetLastError();
    LAUNCH(k_fill<uint64_t>, dim3(fill_blocks((size_t) (tmask + 1), sizeof(uint64_t))), dim3(kBlock), 0, st, table, kEmpty, (size_t) (tmask + 1));
    LAUNCH(k_fill<uint64_t>, dim3(fill_blocks((size_t) 2, sizeof(uint64_t))), dim3(kBlock), 0, st,
           reinterpret_cast<uint64_t *>(bits), (uint64_t) 0, (size_t) 2);
    if (nW_max > 0) LAUNCH(k_stamp, dim3(blocks_for(nW_max)), dim3(kBlock), 0, st, W, nW, stamp, tid);
    LAUNCH(k_path_bits, dim3(blocks_for(nsp_max)), dim3(kBlock), 0, st, spx, spm, mw, nsp, stamp, tid, bits);
    if (nW_max > 0)
        LAUNCH(k_path_insert, dim3(blocks_for(nW_max)), dim3(kBlock), 0, st, D, smask, W, nW, bits, table, tmask);
    LAUNCH(k_path_check, dim3(blocks_for(nsp_max)), dim3(kBlock), 0, st, D, smask, spx, spm, mw, nsp, stamp, tid,
           table, tmask, outx, outm, nout);
    return hipGetLastError();
}

hipError_t launch_redundancy_seed(const int *W, const int *nW_dev, int nW, int *stamp, int tid, const Entry *E,
                                  const int *nE, dmask_t smask, const int *comp, int *list, int *nlist, int *mn,
                                  float r, int *keep, uint8_t *lab, hipStream_t st) {
    LAUNCH(k_select, dim3(1), dim3(kBlock), 0, st, E, nE, smask, list, nlist);
    if (nW > 0) LAUNCH(k_stamp, dim3(blocks_for(nW)), dim3(kBlock), 0, st, W, nW_dev, stamp, tid);
    LAUNCH(k_redundancy, dim3(256), dim3(kBlock), 0, st, stamp, tid, E, list, nlist, mn);
    LAUNCH(k_decide, dim3(1), dim3(64), 0, st, nlist, mn, nW, r, keep);
    if (nW > 0) LAUNCH(k_seed_if, dim3(blocks_for(nW)), dim3(kBlock), 0, st, W, nW, comp, keep, lab);
    return hipGetLastError();
}

hipError_t launch_join(const uint8_t *lab, int n, const int *compS, const int *csizeS, int min_size, float r,
                       Entry *E, const int *nE, dmask_t smask, int eS, uint8_t *inresS, int *list, int *nlist,
                       uint8_t *keepf, uint8_t *drop, uint8_t *fresh, hipStream_t st) {
    LAUNCH(k_join_keep, dim3(blocks_for(n)), dim3(kBlock), 0, st, lab, n, csizeS, min_size, keepf, drop);
    LAUNCH(k_select, dim3(1), dim3(kBlock), 0, st, E, nE, smask, list, nlist);
    LAUNCH(k_join_check, dim3(256), dim3(kBlock), 0, st, compS, keepf, csizeS, r, E, list, nlist, drop);
    LAUNCH(k_join_apply, dim3(blocks_for(n)), dim3(kBlock), 0, st, keepf, drop, n, inresS, fresh, E, eS);
    LAUNCH(k_join_members, dim3(blocks_for(n)), dim3(kBlock), 0, st, compS, n, fresh, E, eS);
    return hipGetLastError();
}

hipError_t launch_set_entry(Entry *E, int idx, const Entry &e, int *nE, int count, hipStream_t st) {
    LAUNCH(k_set_entry, dim3(1), dim3(64), 0, st, E, idx, e, nE, count);
    return hipGetLastError();
}

hipError_t launch_emit(const EmitRow *rows, int r, int n, int *flags, int *idx, void *labels, int width, int *counts,
                       void *scan_temp, size_t scan_bytes, hipStream_t st) {
    if (r <= 0) return hipSuccess;
    const size_t items = (size_t) r * n + 1;
    LAUNCH(k_fill<int>, dim3(fill_blocks((size_t) 1, sizeof(int))), dim3(kBlock), 0, st, flags + (size_t) r * n, 0, (size_t) 1);
    LAUNCH(k_emit_roots, dim3(blocks_for(n), r), dim3(kBlock), 0, st, rows, n, flags);
    hipError_t e = launch_exclusive_sum(scan_temp, scan_bytes, flags, idx, (int) items, st);
    if (e != hipSuccess) return e;
    if (width == 1)
        LAUNCH(k_emit_labels<uint8_t>, dim3(blocks_for(n), r), dim3(kBlock), 0, st, rows, n, idx,
               static_cast<uint8_t *>(labels));
    else if (width == 2)
        LAUNCH(k_emit_labels<uint16_t>, dim3(blocks_for(n), r), dim3(kBlock), 0, st, rows, n, idx,
               static_cast<uint16_t *>(labels));
    else
        LAUNCH(k_emit_labels<int>, dim3(blocks_for(n), r), dim3(kBlock), 0, st, rows, n, idx,
               static_cast<int *>(labels));
    LAUNCH(k_emit_counts, dim3(blocks_for(r)), dim3(kBlock), 0, st, idx, n, r, counts);
    return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t *p, uint8_t v, size_t count, hipStream_t st) {
    if (!count) return hipSuccess;
    LAUNCH(k_fill<uint8_t>, dim3(fill_blocks(count, sizeof(uint8_t))), dim3(kBlock), 0, st, p, v, count);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t v, size_t count, hipStream_t st) {
    if (!count) return hipSuccess;
    LAUNCH(k_fill<uint64_t>, dim3(fill_blocks(count, sizeof(uint64_t))), dim3(kBlock), 0, st, p, v, count);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, int v, size_t count, hipStream_t st) {
    if (!count) return hipSuccess;
    LAUNCH(k_fill<int>, dim3(fill_blocks(count, sizeof(int))), dim3(kBlock), 0, st, p, v, count);
    return hipGetLastError();
}

hipError_t launch_gather(const int *src, const int *idx, int count, int *out, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    LAUNCH(k_gather, dim3(blocks_for(count)), dim3(kBlock), 0, st, src, idx, count, out);
    return hipGetLastError();
}

hipError_t launch_iota(int *p, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    LAUNCH(k_iota, dim3(blocks_for(n)), dim3(kBlock), 0, st, p, n);
    return hipGetLastError();
}

size_t rank_order_temp_bytes(int n) { return sort_temp_bytes(n); }

// Least-significant dimension first, one stable radix sort per dimension (one
// or two 8-bit digit passes for cells below ecells), starting from ascending
// ids: the result is sorted by the cells in `level` order with ties by id.
hipError_t launch_rank_order(const uint16_t *cellE, int n, int d, int ecells, const int *level, int *perm, int *tmp,
                             uint16_t *keys, uint16_t *keys2, void *temp, size_t temp_bytes, hipStream_t st) {
    LAUNCH(k_iota, dim3(blocks_for(n)), dim3(kBlock), 0, st, perm, n);
    int bits = 1;
    while (bits < 16 && (1 << bits) < ecells) ++bits;
    const bool two = bits > 8;
    int *a = perm, *b = tmp;
    for (int t = d - 1; t >= 0; --t) {
        LAUNCH(k_rank_keys, dim3(blocks_for(n)), dim3(kBlock), 0, st, cellE, n, level[t], a, keys);
        // two passes: keys -> keys2 -> keys and a -> b -> a; one pass: keys -> keys2 and a -> b
        hipError_t e = launch_sort_pairs_u16(keys, keys2, two ? keys : keys2, a, b, two ? a : b, n, bits, temp,
                                             temp_bytes, st);
        if (e != hipSuccess) return e;
        if (!two) {
            int *x = a;
            a = b;
            b = x;
        }
    }
    if (a != perm) {
        hipError_t e = hipMemcpyAsync(perm, a, (size_t) n * sizeof(int), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_root_paths(const dmask_t *nearm, int n, int *outx, uint64_t *outm, int mw, int *nout, hipStream_t st) {
    LAUNCH(k_fill<int>, dim3(fill_blocks((size_t) 1, sizeof(int))), dim3(kBlock), 0, st, nout, 0, (size_t) 1);
    LAUNCH(k_root_paths, dim3(blocks_for(n)), dim3(kBlock), 0, st, nearm, n, outx, outm, mw, nout);
    return hipGetLastError();
}

}  // namespace inscy
