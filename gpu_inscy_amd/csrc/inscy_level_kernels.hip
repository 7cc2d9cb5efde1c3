// Level-synchronous passes of the MI355X INSCY engine (gfx950, wave64).
//
// The same per-tree and per-subspace work as the depth-first kernels in
// inscy_kernels.hip, but every launch covers a whole batch: all trees of one
// recursion level (merge flags, partition, path prune) or many subspaces (grid
// index, neighbour counts, dense components).  A block belongs to exactly one
// segment, found by a binary search of its block offsets, so a wave never
// mixes segments and the wave-aggregated compaction stays valid.
//
// Floating point follows inscy_kernels.hip: no contraction, and the distance
// test is sum <= s_star.

#include <algorithm>
#include <climits>

#include <hipcub/hipcub.hpp>

#include "inscy_device.h"
#include "inscy_internal.h"

#pragma clang fp contract(off)

namespace inscy {

namespace {

constexpr int QPB = kBlock / kLanesPerQuery;

// segment owning block `blk`: the last s with blk_off[s] <= blk (empty
// segments share an offset with their successor and are skipped)
__device__ __forceinline__ int seg_of(const int *__restrict__ blk_off, int nseg, int blk) {
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (blk_off[mid] <= blk)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int origin(int x) { return x >= 0 ? x : ~x; }

// ------------------------------------------------------------ merge flags of every tree of a level
// (k_flags_all per tree: ScyTreeNode.cu:837-851 with restrict_node :262-267)
__global__ void k_b_flags(const FlagTree *__restrict__ T, const int *__restrict__ blk_off, int nt,
                          const uint8_t *__restrict__ cellC, int n, int cells, int *__restrict__ flags) {
    const int t = seg_of(blk_off, nt, blockIdx.x);
    const FlagTree F = T[t];
    const int i = (blockIdx.x - blk_off[t]) * blockDim.x + threadIdx.x;
    if (i >= *F.nsp) return;
    const int x = origin(F.spx[i]);
    uint64_t m = F.first >= 64 ? 0ull : (F.spm[i] >> F.first);
    int *fl = flags + F.flag_off;
    while (m) {
        int b = __ffsll((long long) m) - 1;
        m &= m - 1;
        fl[b * cells + cellC[(size_t) (b + F.first) * n + x]] = 1;
    }
}

// ------------------------------------------------------------ grid indexes of a batch of subspaces
__global__ void k_b_keys(DevData D, const BatchSub *__restrict__ subs, int *__restrict__ keys, int *__restrict__ rank,
                         int *__restrict__ hist) {
    const BatchSub &B = subs[blockIdx.y];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.n) return;
    int key = 0;
    for (int a = 0; a < B.G.j; ++a) key = key * B.G.C + D.cellE[(size_t) B.S.dims[B.G.tI[a]] * D.n + p];
    keys[B.base + p] = key;
    rank[B.base + p] = atomicAdd(&hist[B.hoff + key], 1);
}

// `off` is one exclusive scan over the grid cells of every subspace of the
// batch, so positions are global: subspace b holds [b*n, (b+1)*n).  The
// scatter writes one int per point; the coordinates and cells are then
// gathered in sorted order, so their writes are coalesced.
__global__ void k_b_scatter(DevData D, const BatchSub *__restrict__ subs, const int *__restrict__ keys,
                            const int *__restrict__ rank, const int *__restrict__ off, int *__restrict__ order) {
    const BatchSub &B = subs[blockIdx.y];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.n) return;
    order[off[B.hoff + keys[B.base + p]] + rank[B.base + p]] = p;
}

__global__ void k_b_gather(DevData D, const BatchSub *__restrict__ subs, const int *__restrict__ order,
                           float *__restrict__ Ys, uint16_t *__restrict__ Es) {
    const BatchSub &B = subs[blockIdx.y];
    const int n = D.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = order[B.base + i];
    for (int t = 0; t < B.S.k; ++t) {
        const int s = B.S.dims[t];
        Ys[B.coord + (long long) t * n + i] = D.Xt[(size_t) s * n + p];
        Es[B.coord + (long long) t * n + i] = D.cellE[(size_t) s * n + p];
    }
}

// ------------------------------------------------------------ query lists in sorted order
__global__ void k_b_qmark(const BatchSub *__restrict__ subs, const QuerySeg *__restrict__ segs,
                          const int *__restrict__ blk_off, int nseg, const int *__restrict__ keys,
                          const int *__restrict__ rank, const int *__restrict__ off, uint8_t *__restrict__ qflag) {
    const int s = seg_of(blk_off, nseg, blockIdx.x);
    const QuerySeg Q = segs[s];
    const int i = (blockIdx.x - blk_off[s]) * blockDim.x + threadIdx.x;
    if (i >= Q.len) return;
    const BatchSub &B = subs[Q.b];
    const int p = Q.P[i];
    qflag[off[B.hoff + keys[B.base + p]] + rank[B.base + p]] = 1;
}

// qstart[b] = first index of the selected (ascending global) positions that is
// >= b*n; qstart[nsub] = the number selected
__global__ void k_b_qstart(const int *__restrict__ qlist, const int *__restrict__ nsel, int nsub, int n,
                           int *__restrict__ qstart) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nsub) return;
    const int total = *nsel;
    if (b == nsub) {
        qstart[b] = total;
        return;
    }
    const long long key = (long long) b * n;
    int lo = 0, hi = total;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (qlist[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    qstart[b] = lo;
}

// ------------------------------------------------------------ neighbour counts
// k_grid_count over a batch: LIST = the compacted queries of each subspace
// (blocks by segment), else every sorted position of subspace blockIdx.y whose
// count is still missing (components pass, with the subset bound).
template <bool LIST>
__global__ __launch_bounds__(kBlock) void k_b_count(const BatchSub *__restrict__ subs, const int *__restrict__ blk_off,
                                                    int nsub, int n, const int *__restrict__ off,
                                                    const int *__restrict__ order, const float *__restrict__ Ys,
                                                    const uint16_t *__restrict__ Es, const int *__restrict__ qlist,
                                                    const int *__restrict__ qstart, float s_star,
                                                    int *__restrict__ nb, unsigned long long *__restrict__ stat) {
    constexpr int LPQ = kLanesPerQuery;
    extern __shared__ char smem[];
    int b, blk;
    if (LIST) {
        b = seg_of(blk_off, nsub, blockIdx.x);
        blk = blockIdx.x - blk_off[b];
    } else {
        b = blockIdx.y;
        blk = blockIdx.x;
    }
    const BatchSub &B = subs[b];
    const int k = B.S.k;
    float *sy = reinterpret_cast<float *>(smem);
    int *se = reinterpret_cast<int *>(smem + sizeof(float) * QPB * k);
    const int lq = threadIdx.x / LPQ, sub = threadIdx.x % LPQ;
    int i = 0;
    bool active;
    if (LIST) {
        const int q = qstart[b] + blk * QPB + lq;
        active = q < qstart[b + 1];
        if (active) i = (int) (qlist[q] - B.base);
    } else {
        i = blk * QPB + lq;
        active = i < n;
    }
    const float *Yb = Ys + B.coord;
    const uint16_t *Eb = Es + B.coord;
    int *cnt = B.cnt;
    const int p = active ? order[B.base + i] : 0;
    active = active && cnt[p] < 0;
    // N_S(p) is a subset of N_S'(p) for S' a subset of S: the smallest
    // complete count of the subspaces S minus one dimension, when below both
    // thresholds of S, decides p without a scan (and stays an upper bound)
    if (active && B.nbound > 0) {
        int bb = INT_MAX;
        for (int t = 0; t < B.nbound; ++t) {
            const int v = B.bound[t][p];
            if (v >= 0) bb = min(bb, v);
        }
        if (bb < B.lo) {
            if (sub == 0) cnt[p] = bb;
            active = false;
        }
    }
    if (sub == 0 && active) stage_query(B.S, n, Yb, Eb, i, sy + lq, se + lq, QPB);
    __syncthreads();
    long long seen = 0;
    int found = 0;
    if (active) {
        const float *qy = sy + lq;
        const int *qe = se + lq;
        const int gshift = (__lane_id() / LPQ) * LPQ;
        const unsigned long long below = (1ull << sub) - 1ull;
        const int cap = B.cap;
        const int base0 = (int) B.base;
        int *nbp = nb + 2 * (B.base + p);
        for_ranges(B.G, off + B.hoff, qe, QPB, [&](int a, int e) {
            a -= base0;
            e -= base0;
            for (int base = a; base < e && found < cap; base += LPQ) {
                const int jj = base + sub;
                const bool hit = jj < e && jj != i && is_neighbour(B.S, B.G, n, Yb, Eb, qy, qe, QPB, jj, s_star);
                const unsigned long long mine = (__ballot(hit) >> gshift) & ((1ull << LPQ) - 1ull);
                // the first two neighbours found: the sample edges of the components pass
                if (hit && found < 2) {
                    const int at = found + __popcll(mine & below);
                    if (at < 2) nbp[at] = order[B.base + jj];
                }
                found += __popcll(mine);
                seen += min(LPQ, e - base);
            }
        });
        if (sub == 0) cnt[p] = found;
    }
    if (stat) {
        // per wave into one of 64 slot groups: [0] candidates examined, [1]
        // points counted, [2] algorithmic bytes (4k + 4 per point, 4 + 4k per
        // neighbour established)
        const bool mine = sub == 0 && active;
        long long done = mine ? 1 : 0;
        long long abytes = mine ? (4ll * k + 4) + (long long) found * (4 + 4ll * k) : 0;
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

// ------------------------------------------------------------ dense components of a batch
// (k_dense_init .. k_uf_flatten per subspace; grid.y = subspace)
__global__ void k_b_dense_init(const BatchSub *__restrict__ subs, int n, const int *__restrict__ order,
                               uint8_t *__restrict__ dsorted, int *__restrict__ par) {
    const BatchSub &B = subs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = order[B.base + i];
    const bool dense = B.cnt[p] >= B.thr_dense;
    dsorted[B.base + i] = dense ? 1 : 0;
    par[B.base + p] = dense ? p : -1;
    B.csize[p] = 0;
}

// Sample phase of the components: every dense point links to the dense ones
// among the first two neighbours its count pass recorded (any edges would do:
// the full phase below links every point outside the giant component to all
// its dense neighbours, so the result does not depend on the sample).
__global__ void k_b_link_sample(const BatchSub *__restrict__ subs, int n, const int *__restrict__ nb,
                                int *__restrict__ par) {
    const BatchSub &B = subs[blockIdx.y];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int *cnt = B.cnt;
    const int c = cnt[p];
    if (c < B.thr_dense) return;
    int *pr = par + B.base;
    const int *nbp = nb + 2 * (B.base + p);
    const int k = min(c, 2);
    for (int t = 0; t < k; ++t) {
        const int q = nbp[t];
        if (q >= 0 && q < n && cnt[q] >= B.thr_dense) uf_union(pr, p, q);
    }
}

__global__ void k_b_uf_compress(const BatchSub *__restrict__ subs, int n, int *__restrict__ par) {
    int *pr = par + subs[blockIdx.y].base;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ld_relaxed(&pr[i]) >= 0) st_relaxed(&pr[i], uf_find(pr, i));
}

// one block of 256 per subspace: the most frequent root among 256 evenly
// spaced points
__global__ void k_b_giant(const BatchSub *__restrict__ subs, int n, const int *__restrict__ par,
                          int *__restrict__ giant) {
    __shared__ int roots[256];
    __shared__ int best_c[256];
    __shared__ int best_r[256];
    const int *pr = par + subs[blockIdx.x].base;
    const int t = threadIdx.x;
    const int p = (int) (((long long) t * n) / 256);
    roots[t] = p < n ? ld_relaxed(&pr[p]) : -1;
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
    if (t == 0) giant[blockIdx.x] = best_c[0] > 0 ? best_r[0] : -1;
}

__global__ __launch_bounds__(kBlock) void k_b_link_full(const BatchSub *__restrict__ subs, int n,
                                                        const int *__restrict__ off, const int *__restrict__ order,
                                                        const float *__restrict__ Ys, const uint16_t *__restrict__ Es,
                                                        const uint8_t *__restrict__ dsorted, float s_star,
                                                        const int *__restrict__ giant, int *__restrict__ par) {
    constexpr int LPQ = kLanesPerQuery;
    extern __shared__ char smem[];
    const BatchSub &B = subs[blockIdx.y];
    const int k = B.S.k;
    float *sy = reinterpret_cast<float *>(smem);
    int *se = reinterpret_cast<int *>(smem + sizeof(float) * QPB * k);
    const int lq = threadIdx.x / LPQ, sub = threadIdx.x % LPQ;
    const int i = blockIdx.x * QPB + lq;
    const uint8_t *ds = dsorted + B.base;
    const int *ord = order + B.base;
    int *pr = par + B.base;
    const float *Yb = Ys + B.coord;
    const uint16_t *Eb = Es + B.coord;
    bool active = i < n && ds[i];
    const int p = active ? ord[i] : 0;
    if (active && ld_relaxed(&pr[p]) == giant[blockIdx.y]) active = false;
    if (sub == 0 && active) stage_query(B.S, n, Yb, Eb, i, sy + lq, se + lq, QPB);
    __syncthreads();
    if (!active) return;
    const float *qy = sy + lq;
    const int *qe = se + lq;
    const int base0 = (int) B.base;
    int rp = uf_find(pr, p);
    for_ranges(B.G, off + B.hoff, qe, QPB, [&](int a, int e) {
        a -= base0;
        e -= base0;
        for (int jj = a + sub; jj < e; jj += LPQ) {
            if (jj == i || !ds[jj] || !is_neighbour(B.S, B.G, n, Yb, Eb, qy, qe, QPB, jj, s_star)) continue;
            const int q = ord[jj];
            if (uf_find(pr, q) == rp) continue;
            uf_union(pr, p, q);
            rp = uf_find(pr, p);
        }
    });
}

__global__ void k_b_uf_flatten(const BatchSub *__restrict__ subs, int n, int *__restrict__ par) {
    const BatchSub &B = subs[blockIdx.y];
    int *pr = par + B.base;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1;
    if (i < n && ld_relaxed(&pr[i]) >= 0) c = uf_find(pr, i);
    wave_add_by_key(B.csize, c, c >= 0);
    if (i < n) B.comp[i] = c;
}

// ------------------------------------------------------------ partition of every (tree, dimension) pair
// k_part_count / k_part_scatter per pair.  Only the weak-dense points and the
// paths are written; a path's origin keeps its "point of this tree" bit when
// the origin survives the weak-density prune of the group.
__global__ void k_b_part_count(const PartPair *__restrict__ pairs, const int *__restrict__ blk_off, int np,
                               const uint8_t *__restrict__ cellC, int n, int thr) {
    __shared__ int h[3 * 256];
    const int j = seg_of(blk_off, np, blockIdx.x);
    const PartPair &A = pairs[j];
    const int blk = blockIdx.x - blk_off[j];
    const int G = A.G;
    const uint8_t *cm = cellC + (size_t) A.m * n;
    for (int t = threadIdx.x; t < 3 * G; t += blockDim.x) h[t] = 0;
    __syncthreads();
    if (blk < A.bp) {
        const int i = blk * blockDim.x + threadIdx.x;
        if (i < A.nP) {
            const int p = A.P[i];
            const int g = A.gid[cm[p]];
            atomicAdd(&h[g], 1);
            if (A.weakcnt[p] >= thr) atomicAdd(&h[G + g], 1);
        }
    } else {
        const int i = (blk - A.bp) * blockDim.x + threadIdx.x;
        if (i < A.nsp) {
            const uint64_t mk = A.m >= 63 ? 0ull : (A.spm[i] & (~0ull << (A.m + 1)));
            if (mk) atomicAdd(&h[2 * G + A.gid[cm[origin(A.spx[i])]]], 1);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * G; t += blockDim.x)
        if (h[t]) atomicAdd(&A.cnt[t], h[t]);
}

__global__ void k_b_part_scatter(const PartPair *__restrict__ pairs, const int *__restrict__ blk_off, int np,
                                 const uint8_t *__restrict__ cellC, int n, int thr) {
    __shared__ int base[3 * 256];
    const int j = seg_of(blk_off, np, blockIdx.x);
    const PartPair &A = pairs[j];
    const int blk = blockIdx.x - blk_off[j];
    const int G = A.G;
    const uint8_t *cm = cellC + (size_t) A.m * n;
    if (threadIdx.x < 3) {
        int acc = 0;
        for (int g = 0; g < G; ++g) {
            base[threadIdx.x * G + g] = acc;
            acc += A.cnt[threadIdx.x * G + g];
        }
    }
    __syncthreads();
    int *cur = A.cnt + 3 * G;
    if (blk < A.bp) {
        const int i = blk * blockDim.x + threadIdx.x;
        const bool in = i < A.nP;
        const int p = in ? A.P[i] : 0;
        const int g = in ? A.gid[cm[p]] : 0;
        const bool weak = in && A.weakcnt[p] >= thr;
        const int sw = wave_add_by_key(cur + G, g, weak);
        if (weak) A.outW[base[G + g] + sw] = p;
    } else {
        const int i = (blk - A.bp) * blockDim.x + threadIdx.x;
        uint64_t mk = 0;
        int x = 0, g = 0;
        if (i < A.nsp) {
            mk = A.m >= 63 ? 0ull : (A.spm[i] & (~0ull << (A.m + 1)));
            const int xr = A.spx[i];
            const int xo = origin(xr);
            g = A.gid[cm[xo]];
            x = (xr >= 0 && A.weakcnt[xo] >= thr) ? xo : ~xo;
        }
        const int sx = wave_add_by_key(cur + 2 * G, g, mk != 0);
        if (mk) {
            const int pos = base[2 * G + g] + sx;
            A.outx[pos] = x;
            A.outm[pos] = mk;
        }
    }
}

// ------------------------------------------------------------ path prune of every surviving tree
// (k_path_bits / k_path_insert / k_path_check per tree; "origin in the tree"
// is the sign of the stored origin instead of a stamp)
__global__ void k_b_path_bits(const PruneChild *__restrict__ C, const int *__restrict__ blk_p, int nc) {
    const int c = seg_of(blk_p, nc, blockIdx.x);
    const PruneChild &P = C[c];
    const int i = (blockIdx.x - blk_p[c]) * blockDim.x + threadIdx.x;
    uint64_t m = 0;
    if (i < P.nsp && P.spx[i] < 0) m = P.spm[i];
    for (int o = 32; o > 0; o >>= 1) m |= __shfl_xor(m, o);
    if (__lane_id() == 0 && m) atomicOr(P.bits, (unsigned long long) m);
}

__global__ void k_b_path_insert(DevData D, const PruneChild *__restrict__ C, const int *__restrict__ blk_w, int nc) {
    const int c = seg_of(blk_w, nc, blockIdx.x);
    const PruneChild &P = C[c];
    const int i = (blockIdx.x - blk_w[c]) * blockDim.x + threadIdx.x;
    if (i >= P.nW) return;
    const uint64_t need = *P.bits;
    if (!need) return;
    const int w = P.W[i];
    const uint64_t smask = P.smask;
    uint64_t h = 0x243F6A8885A308D3ull;
    for (int k = 0; k < D.d; ++k) {
        if ((smask >> k) & 1ull) continue;
        h = mix64(h + D.cellC[(size_t) k * D.n + w] + 1);
        if (!((need >> k) & 1ull)) continue;
        const uint64_t e = mix64(h ^ (0x9E3779B97F4A7C15ull * (uint64_t) (k + 1)));
        const uint64_t entry = (e & 0xFFFFFFFF00000000ull) | (uint32_t) w;
        uint64_t slot = e & P.tmask;
        while (true) {
            uint64_t cur = __hip_atomic_load(&P.table[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == kEmpty) {
                cur = atomicCAS((unsigned long long *) &P.table[slot], (unsigned long long) kEmpty,
                                (unsigned long long) entry);
                if (cur == kEmpty) break;
            }
            if ((cur >> 32) == (entry >> 32) && same_prefix(D, smask, w, (int) (uint32_t) cur, k)) break;
            slot = (slot + 1) & P.tmask;
        }
    }
}

__global__ void k_b_path_check(DevData D, const PruneChild *__restrict__ C, const int *__restrict__ blk_p, int nc) {
    const int c = seg_of(blk_p, nc, blockIdx.x);
    const PruneChild &P = C[c];
    const int i = (blockIdx.x - blk_p[c]) * blockDim.x + threadIdx.x;
    int x = 0;
    uint64_t m = 0;
    if (i < P.nsp) {
        x = P.spx[i];
        m = P.spm[i];
        if (x < 0) {
            const int xo = ~x;
            const uint64_t smask = P.smask;
            uint64_t h = 0x243F6A8885A308D3ull;
            uint64_t keep = 0;
            for (int k = 0; k < D.d; ++k) {
                if ((smask >> k) & 1ull) continue;
                h = mix64(h + D.cellC[(size_t) k * D.n + xo] + 1);
                if (!((m >> k) & 1ull)) continue;
                const uint64_t e = mix64(h ^ (0x9E3779B97F4A7C15ull * (uint64_t) (k + 1)));
                uint64_t slot = e & P.tmask;
                while (true) {
                    const uint64_t cur =
                        __hip_atomic_load(&P.table[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (cur == kEmpty) break;
                    if ((cur >> 32) == (e >> 32) && same_prefix(D, smask, xo, (int) (uint32_t) cur, k)) {
                        keep |= 1ull << k;
                        break;
                    }
                    slot = (slot + 1) & P.tmask;
                }
            }
            m = keep;
        }
    }
    const int slot = wave_slot(m != 0, P.nout);
    if (m != 0) {
        P.outx[slot] = x;
        P.outm[slot] = m;
    }
}

// ------------------------------------------------------------ replay of pruneRedundancy and join
// The depth-first order of the reference is replayed with one launch per
// surviving tree and four per join.  `list` holds the result-table rows whose
// subspace contains S (fixed once the expansion is done); a row counts while
// its nkept > 0, which the joins earlier in stream order have set.  A tree is
// known by its index; tstamp[p] is the tree of S whose weak-dense points hold p.

// pruneRedundancy (ScyTreeNode.cu:669-734): per listed superspace H in the
// result (one block per H), the smallest result cluster of H holding a point of
// the tree; the largest of these minima decides.  The last block to finish
// writes the decision.
__global__ void k_rep_redundancy(const Entry *__restrict__ E, const int *__restrict__ list, int nlist,
                                 const int *__restrict__ tstamp, int tree, int nW, float r, int *__restrict__ mx,
                                 unsigned *__restrict__ done, int *__restrict__ keepv) {
    __shared__ int wmin[kBlock / 64];
    __shared__ bool last;
    for (int h = blockIdx.x; h < nlist; h += gridDim.x) {
        const Entry e = E[list[h]];
        if (e.nkept <= 0) continue;
        int best = INT_MAX;
        for (int t = threadIdx.x; t < e.nmem; t += blockDim.x) {
            const int q = e.mem[t];
            if (tstamp[q] == tree) best = min(best, e.csize[e.comp[q]]);
        }
        for (int o = 32; o > 0; o >>= 1) best = min(best, __shfl_down(best, o));
        if (__lane_id() == 0) wmin[threadIdx.x / 64] = best;
        __syncthreads();
        if (threadIdx.x == 0) {
            int b = INT_MAX;
            for (int w = 0; w < kBlock / 64; ++w) b = min(b, wmin[w]);
            if (b != INT_MAX) atomicMax(mx, b);
        }
        __syncthreads();
    }
    __threadfence();
    if (threadIdx.x == 0) last = atomicAdd(done, 1u) == gridDim.x - 1;
    __syncthreads();
    if (last && threadIdx.x == 0) {
        const int m = __hip_atomic_load(mx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const float lhs = (float) nW * r;
        keepv[tree] = (m == 0 || (double) lhs > (double) m) ? 1 : 0;
        __hip_atomic_store(mx, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// join (util.cu:655-743) of the clusters of the kept trees [t_lo, t_hi) of one
// (parent, dimension) pair: cluster (ClusteringCpu.cu:317-389) labels the dense
// component of every point of a kept tree; components of at least min_size
// are candidates
__global__ void k_rep_join_mark(const int *__restrict__ tstamp, int t_lo, int t_hi, const int *__restrict__ keepv,
                                const int *__restrict__ comp, const int *__restrict__ csize, int min_size, int n,
                                uint8_t *__restrict__ keepf) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int t = tstamp[p];
    if (t < t_lo || t >= t_hi || !keepv[t]) return;
    const int c = comp[p];
    if (c >= 0 && csize[c] >= min_size) keepf[c] = 1;
}

// a candidate K is dropped when one of its points lies in a result cluster of
// a superspace H larger than |K| * r (one block per H)
__global__ void k_rep_join_check(const Entry *__restrict__ E, const int *__restrict__ list, int nlist,
                                 const int *__restrict__ compS, const uint8_t *__restrict__ keepf,
                                 const int *__restrict__ csizeS, float r, uint8_t *__restrict__ drop) {
    for (int h = blockIdx.x; h < nlist; h += gridDim.x) {
        const Entry e = E[list[h]];
        if (e.nkept <= 0) continue;
        for (int t = threadIdx.x; t < e.nmem; t += blockDim.x) {
            const int q = e.mem[t];
            const int K = compS[q];
            if (K < 0 || !keepf[K]) continue;
            if ((float) csizeS[K] * r < (float) e.csize[e.comp[q]]) drop[K] = 1;
        }
    }
}

// survivors join the result of S; the candidate flags are cleared for the next join
__global__ void k_rep_join_apply(uint8_t *__restrict__ keepf, uint8_t *__restrict__ drop, int n,
                                 uint8_t *__restrict__ inres, uint8_t *__restrict__ fresh, Entry *__restrict__ E,
                                 int eS) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    int added = 0;
    if (k < n) {
        uint8_t f = 0;
        if (keepf[k] && !drop[k] && !inres[k]) {
            inres[k] = 1;
            f = 1;
            added = 1;
        }
        fresh[k] = f;
        keepf[k] = 0;
        drop[k] = 0;
    }
    for (int o = 32; o > 0; o >>= 1) added += __shfl_down(added, o);
    if (__lane_id() == 0 && added) atomicAdd(&E[eS].nkept, added);
}

// members of the newly kept clusters join S's member list
__global__ void k_rep_join_members(const int *__restrict__ compS, int n, const uint8_t *__restrict__ fresh,
                                   Entry *__restrict__ E, int eS) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (i < n) {
        const int K = compS[i];
        in = K >= 0 && fresh[K];
    }
    const int slot = wave_slot(in, &E[eS].nmem);
    if (in) E[eS].mem[slot] = i;
}

__global__ void k_b_stamp(const CopySeg *__restrict__ segs, const int *__restrict__ blk_off, int ns) {
    const int s = seg_of(blk_off, ns, blockIdx.x);
    const CopySeg S = segs[s];
    const int i = (blockIdx.x - blk_off[s]) * blockDim.x + threadIdx.x;
    if (i < S.len) S.dst[S.src[i]] = S.pad;
}

// ------------------------------------------------------------ small helpers
__global__ void k_b_copy(const CopySeg *__restrict__ segs, const int *__restrict__ blk_off, int ns) {
    const int s = seg_of(blk_off, ns, blockIdx.x);
    const CopySeg S = segs[s];
    const int i = (blockIdx.x - blk_off[s]) * blockDim.x + threadIdx.x;
    if (i < S.len) S.dst[i] = S.src[i];
}

__global__ void k_b_entries(Entry *__restrict__ E, int at, const Entry *__restrict__ src, int count,
                            int *__restrict__ nE) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) E[at + i] = src[i];
    if (threadIdx.x == 0) *nE = at + count;
}

template <class T>
__global__ void k_b_fill(T *p, T v, size_t count) {
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

}  // namespace

#define LAUNCH(kernel, grid, block, lds, st, ...) hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__)

static_assert(kBlockThreads == kBlock && kQueriesPerBlock == QPB, "launch geometry constants disagree");

static unsigned query_blocks_n(int n) { return (unsigned) ((n + QPB - 1) / QPB); }
static size_t count_lds(int kmax) { return (sizeof(float) + sizeof(int)) * QPB * (size_t) (kmax > 0 ? kmax : 1); }

hipError_t launch_b_flags(const FlagTree *T, const int *blk_off, int nt, int nblk, const uint8_t *cellC, int n,
                          int cells, int *flags, hipStream_t st) {
    if (nblk <= 0) return hipSuccess;
    LAUNCH(k_b_flags, dim3(nblk), dim3(kBlock), 0, st, T, blk_off, nt, cellC, n, cells, flags);
    return hipGetLastError();
}

hipError_t launch_b_index(const DevData &D, const BatchSub *subs, int nsub, const BatchIndex &X, hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    LAUNCH(k_b_fill<int>, dim3(blocks_for(X.nhist)), dim3(kBlock), 0, st, X.hist, 0, (size_t) X.nhist);
    LAUNCH(k_b_keys, dim3(blocks_for(D.n), nsub), dim3(kBlock), 0, st, D, subs, X.keys, X.rank, X.hist);
    size_t tb = X.scan_bytes;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(X.scan_temp, tb, X.hist, X.off, X.nhist, st);
    if (e != hipSuccess) return e;
    LAUNCH(k_b_scatter, dim3(blocks_for(D.n), nsub), dim3(kBlock), 0, st, D, subs, X.keys, X.rank, X.off, X.order);
    LAUNCH(k_b_gather, dim3(blocks_for(D.n), nsub), dim3(kBlock), 0, st, D, subs, X.order, X.Ys, X.Es);
    return hipGetLastError();
}

size_t select_temp_bytes(int items) {
    size_t bytes = 0;
    if (hipcub::DeviceSelect::Flagged(nullptr, bytes, hipcub::CountingInputIterator<int>(0), (const uint8_t *) nullptr,
                                      (int *) nullptr, (int *) nullptr, items) != hipSuccess)
        return 1u << 20;
    return bytes;
}

hipError_t launch_b_queries(const BatchSub *subs, int nsub, int n, const QuerySeg *segs, const int *blk_off,
                            int nseg, int nblk, const BatchIndex &X, hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    const size_t nb = (size_t) nsub * n;
    LAUNCH(k_b_fill<uint8_t>, dim3(blocks_for(nb)), dim3(kBlock), 0, st, X.qflag, (uint8_t) 0, nb);
    if (nblk > 0)
        LAUNCH(k_b_qmark, dim3(nblk), dim3(kBlock), 0, st, subs, segs, blk_off, nseg, X.keys, X.rank, X.off, X.qflag);
    size_t tb = X.select_bytes;
    hipError_t e = hipcub::DeviceSelect::Flagged(X.select_temp, tb, hipcub::CountingInputIterator<int>(0), X.qflag,
                                                 X.qlist, X.nsel, (int) nb, st);
    if (e != hipSuccess) return e;
    LAUNCH(k_b_qstart, dim3(blocks_for(nsub + 1)), dim3(kBlock), 0, st, X.qlist, X.nsel, nsub, n, X.qstart);
    return hipGetLastError();
}

hipError_t launch_b_count(const BatchSub *subs, int nsub, const int *blk_off, int nblk, int n, int kmax,
                          const BatchIndex &X, float s_star, unsigned long long *stat, const KernelTimer &tm,
                          hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    if (blk_off && nblk <= 0) return hipSuccess;
    if (tm.start) (void) hipEventRecord(tm.start, st);
    if (blk_off)
        LAUNCH(k_b_count<true>, dim3(nblk), dim3(kBlock), count_lds(kmax), st, subs, blk_off, nsub, n, X.off, X.order,
               X.Ys, X.Es, X.qlist, X.qstart, s_star, X.nb, stat);
    else
        LAUNCH(k_b_count<false>, dim3(query_blocks_n(n), nsub), dim3(kBlock), count_lds(kmax), st, subs,
               (const int *) nullptr, nsub, n, X.off, X.order, X.Ys, X.Es, X.qlist, X.qstart, s_star, X.nb, stat);
    if (tm.stop) (void) hipEventRecord(tm.stop, st);
    return hipGetLastError();
}

hipError_t launch_b_components(const BatchSub *subs, int nsub, int n, int kmax, const BatchIndex &X, float s_star,
                               hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    LAUNCH(k_b_dense_init, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.order, X.dsorted, X.par);
    LAUNCH(k_b_link_sample, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.nb, X.par);
    LAUNCH(k_b_uf_compress, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.par);
    LAUNCH(k_b_giant, dim3(nsub), dim3(256), 0, st, subs, n, X.par, X.giant);
    LAUNCH(k_b_link_full, dim3(query_blocks_n(n), nsub), dim3(kBlock), count_lds(kmax), st, subs, n, X.off, X.order,
           X.Ys, X.Es, X.dsorted, s_star, X.giant, X.par);
    LAUNCH(k_b_uf_flatten, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.par);
    return hipGetLastError();
}

hipError_t launch_b_partition(const PartPair *pairs, const int *blk_off, int npairs, int nblk, const uint8_t *cellC,
                              int n, int thr, hipStream_t st) {
    if (nblk <= 0) return hipSuccess;
    LAUNCH(k_b_part_count, dim3(nblk), dim3(kBlock), 0, st, pairs, blk_off, npairs, cellC, n, thr);
    LAUNCH(k_b_part_scatter, dim3(nblk), dim3(kBlock), 0, st, pairs, blk_off, npairs, cellC, n, thr);
    return hipGetLastError();
}

hipError_t launch_b_prune(const DevData &D, const PruneChild *C, int nc, const int *blk_w, int nblk_w,
                          const int *blk_p, int nblk_p, hipStream_t st) {
    if (nc <= 0 || nblk_p <= 0) return hipSuccess;
    LAUNCH(k_b_path_bits, dim3(nblk_p), dim3(kBlock), 0, st, C, blk_p, nc);
    if (nblk_w > 0) LAUNCH(k_b_path_insert, dim3(nblk_w), dim3(kBlock), 0, st, D, C, blk_w, nc);
    LAUNCH(k_b_path_check, dim3(nblk_p), dim3(kBlock), 0, st, D, C, blk_p, nc);
    return hipGetLastError();
}

hipError_t launch_b_copy(const CopySeg *segs, const int *blk_off, int nseg, int nblk, hipStream_t st) {
    if (nblk <= 0) return hipSuccess;
    LAUNCH(k_b_copy, dim3(nblk), dim3(kBlock), 0, st, segs, blk_off, nseg);
    return hipGetLastError();
}

hipError_t launch_b_stamp(const CopySeg *segs, const int *blk_off, int nseg, int nblk, hipStream_t st) {
    if (nblk <= 0) return hipSuccess;
    LAUNCH(k_b_stamp, dim3(nblk), dim3(kBlock), 0, st, segs, blk_off, nseg);
    return hipGetLastError();
}

hipError_t launch_rep_redundancy(const Entry *E, const int *list, int nlist, const int *tstamp, int tree, int nW,
                                 float r, int *mx, unsigned *done, int *keepv, hipStream_t st) {
    const int grid = std::max(1, std::min(nlist, 256));
    LAUNCH(k_rep_redundancy, dim3(grid), dim3(kBlock), 0, st, E, list, nlist, tstamp, tree, nW, r, mx, done, keepv);
    return hipGetLastError();
}

hipError_t launch_rep_join(const int *tstamp, int t_lo, int t_hi, const int *keepv, const int *compS,
                           const int *csizeS, int min_size, float r, int n, Entry *E, const int *list, int nlist,
                           int eS, uint8_t *inresS, uint8_t *keepf, uint8_t *drop, uint8_t *fresh, hipStream_t st) {
    LAUNCH(k_rep_join_mark, dim3(blocks_for(n)), dim3(kBlock), 0, st, tstamp, t_lo, t_hi, keepv, compS, csizeS,
           min_size, n, keepf);
    if (nlist > 0)
        LAUNCH(k_rep_join_check, dim3(std::min(nlist, 256)), dim3(kBlock), 0, st, E, list, nlist, compS, keepf, csizeS,
               r, drop);
    LAUNCH(k_rep_join_apply, dim3(blocks_for(n)), dim3(kBlock), 0, st, keepf, drop, n, inresS, fresh, E, eS);
    LAUNCH(k_rep_join_members, dim3(blocks_for(n)), dim3(kBlock), 0, st, compS, n, fresh, E, eS);
    return hipGetLastError();
}

hipError_t launch_b_entries(Entry *E, int at, const Entry *src, int count, int *nE, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    LAUNCH(k_b_entries, dim3(1), dim3(kBlock), 0, st, E, at, src, count, nE);
    return hipGetLastError();
}

}  // namespace inscy
