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
// batch, so positions are global: subspace b holds [b*n, (b+1)*n).
__global__ void k_b_scatter(DevData D, const BatchSub *__restrict__ subs, const int *__restrict__ keys,
                            const int *__restrict__ rank, const int *__restrict__ off, int *__restrict__ order,
                            float *__restrict__ Ys, uint16_t *__restrict__ Es) {
    const BatchSub &B = subs[blockIdx.y];
    const int n = D.n;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int pos = off[B.hoff + keys[B.base + p]] + rank[B.base + p];
    order[pos] = p;
    const long long local = pos - B.base;
    for (int t = 0; t < B.S.k; ++t) {
        const int s = B.S.dims[t];
        Ys[B.coord + (long long) t * n + local] = D.Xt[(size_t) s * n + p];
        Es[B.coord + (long long) t * n + local] = D.cellE[(size_t) s * n + p];
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

__global__ void k_b_qcompact(const BatchSub *__restrict__ subs, int n, const uint8_t *__restrict__ qflag,
                             int *__restrict__ qlist, int *__restrict__ qcount) {
    const int b = blockIdx.y;
    const BatchSub &B = subs[b];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool q = i < n && qflag[B.base + i];
    const int slot = wave_slot(q, &qcount[b]);
    if (q && slot < B.qmax) qlist[B.qoff + slot] = i;
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
                                                    const int *__restrict__ qcount, float s_star,
                                                    unsigned long long *__restrict__ stat) {
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
        const int q = blk * QPB + lq;
        active = q < min(qcount[b], B.qmax);
        if (active) i = qlist[B.qoff + q];
    } else {
        i = blk * QPB + lq;
        active = i < n;
    }
    const float *Yb = Ys + B.coord;
    const uint16_t *Eb = Es + B.coord;
    int *cnt = B.cnt;
    const int p = active ? order[B.base + i] : 0;
    active = active && cnt[p] < 0;
    // N_S(p) is a subset of N_S'(p) for S' a subset of S: a complete count of
    // a subspace of S below both thresholds of S decides p without a scan
    if (active && B.bound) {
        const int bb = B.bound[p];
        if (bb >= 0 && bb < B.lo) {
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
        const int cap = B.cap;
        const int base0 = (int) B.base;
        for_ranges(B.G, off + B.hoff, qe, QPB, [&](int a, int e) {
            a -= base0;
            e -= base0;
            for (int base = a; base < e && found < cap; base += LPQ) {
                const int jj = base + sub;
                const bool hit = jj < e && jj != i && is_neighbour(B.S, B.G, n, Yb, Eb, qy, qe, QPB, jj, s_star);
                found += __popcll((__ballot(hit) >> gshift) & ((1ull << LPQ) - 1ull));
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

__global__ __launch_bounds__(kBlock) void k_b_link_sample(const BatchSub *__restrict__ subs, int n,
                                                          const int *__restrict__ off, const int *__restrict__ order,
                                                          const float *__restrict__ Ys, const uint16_t *__restrict__ Es,
                                                          const uint8_t *__restrict__ dsorted, float s_star,
                                                          int *__restrict__ par) {
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
    const bool active = i < n && ds[i];
    if (sub == 0 && active) stage_query(B.S, n, Yb, Eb, i, sy + lq, se + lq, QPB);
    __syncthreads();
    if (!active) return;
    const float *qy = sy + lq;
    const int *qe = se + lq;
    const int p = ord[i];
    const int gshift = (__lane_id() / LPQ) * LPQ;
    const unsigned long long below = (1ull << sub) - 1ull;
    const int base0 = (int) B.base;
    int found = 0;
    for_ranges(B.G, off + B.hoff, qe, QPB, [&](int a, int e) {
        a -= base0;
        e -= base0;
        for (int base = a; base < e && found < 2; base += LPQ) {
            const int jj = base + sub;
            const bool hit = jj < e && jj != i && ds[jj] && is_neighbour(B.S, B.G, n, Yb, Eb, qy, qe, QPB, jj, s_star);
            const unsigned long long mine = (__ballot(hit) >> gshift) & ((1ull << LPQ) - 1ull);
            if (hit && __popcll(mine & below) < 2 - found) uf_union(pr, p, ord[jj]);
            found += min(__popcll(mine), 2 - found);
        }
    });
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
    LAUNCH(k_b_scatter, dim3(blocks_for(D.n), nsub), dim3(kBlock), 0, st, D, subs, X.keys, X.rank, X.off, X.order,
           X.Ys, X.Es);
    return hipGetLastError();
}

hipError_t launch_b_queries(const BatchSub *subs, int nsub, int n, const QuerySeg *segs, const int *blk_off,
                            int nseg, int nblk, const BatchIndex &X, hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    const size_t nb = (size_t) nsub * n;
    LAUNCH(k_b_fill<uint8_t>, dim3(blocks_for(nb)), dim3(kBlock), 0, st, X.qflag, (uint8_t) 0, nb);
    LAUNCH(k_b_fill<int>, dim3(blocks_for(nsub)), dim3(kBlock), 0, st, X.qcount, 0, (size_t) nsub);
    if (nblk > 0)
        LAUNCH(k_b_qmark, dim3(nblk), dim3(kBlock), 0, st, subs, segs, blk_off, nseg, X.keys, X.rank, X.off, X.qflag);
    LAUNCH(k_b_qcompact, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.qflag, X.qlist, X.qcount);
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
               X.Ys, X.Es, X.qlist, X.qcount, s_star, stat);
    else
        LAUNCH(k_b_count<false>, dim3(query_blocks_n(n), nsub), dim3(kBlock), count_lds(kmax), st, subs,
               (const int *) nullptr, nsub, n, X.off, X.order, X.Ys, X.Es, X.qlist, X.qcount, s_star, stat);
    if (tm.stop) (void) hipEventRecord(tm.stop, st);
    return hipGetLastError();
}

hipError_t launch_b_components(const BatchSub *subs, int nsub, int n, int kmax, const BatchIndex &X, float s_star,
                               hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    LAUNCH(k_b_dense_init, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.order, X.dsorted, X.par);
    LAUNCH(k_b_link_sample, dim3(query_blocks_n(n), nsub), dim3(kBlock), count_lds(kmax), st, subs, n, X.off,
           X.order, X.Ys, X.Es, X.dsorted, s_star, X.par);
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

hipError_t launch_b_entries(Entry *E, int at, const Entry *src, int count, int *nE, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    LAUNCH(k_b_entries, dim3(1), dim3(kBlock), 0, st, E, at, src, count, nE);
    return hipGetLastError();
}

}  // namespace inscy
