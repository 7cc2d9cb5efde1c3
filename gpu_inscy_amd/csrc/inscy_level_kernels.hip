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


#include "inscy_device.h"
#include "inscy_internal.h"

#pragma clang fp contract(off)

namespace inscy {

namespace {

constexpr int kBLanes = 4;  // lanes per query of the batched neighbour passes (one quad: the walks sum over it with DPP)
constexpr int QPB = kBlock / kBLanes;
static_assert(kBLanes == 4, "the walks' tally sums a query's lanes as one DPP quad");
static_assert(QPB == kQueriesPerBlock, "queries per block of the host and the kernels");

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
    dmask_t m = mload(F.spm, i, F.mw) & dabove(F.first - 1);
    int *fl = flags + F.flag_off;
    while (m) {
        const int b = dfirst(m);
        m &= m - 1;
        fl[(b - F.first) * cells + cellC[(size_t) b * n + x]] = 1;
    }
}

// ------------------------------------------------------------ grid indexes of a batch of subspaces
// column t of subspace B's sorted coordinates / cells
__device__ __forceinline__ long long col_of(const BatchSub &B, int t, int n) {
    return t == B.pos_m ? B.mcoord : B.coord + (long long) t * n;
}

// grid keys and in-cell ranks of the subspaces that build an index (grid.y =
// position in X.reps); one counting sort per such subspace
__global__ void k_b_keys(DevData D, const BatchSub *__restrict__ subs, const int *__restrict__ reps,
                         int *__restrict__ keys, int *__restrict__ rank, int *__restrict__ hist) {
    const BatchSub &B = subs[reps[blockIdx.y]];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.n) return;
    int key = 0;
    const int jl = B.G.j - 1;
    for (int a = 0; a < jl; ++a) key = key * B.G.C + D.cellE[(size_t) B.S.dims[B.G.tI[a]] * D.n + p];
    // the last index dimension: its cell, or its bucket within the cell
    const int dl = B.S.dims[B.G.tI[jl]];
    if (B.G.F > 1)
        key = key * (B.G.C * B.G.F) + fine_key(D.Xt[(size_t) dl * D.n + p], B.G.C, B.G.F);
    else
        key = key * B.G.C + D.cellE[(size_t) dl * D.n + p];
    keys[B.obase + p] = key;
    rank[B.obase + p] = atomicAdd(&hist[B.hoff + key], 1);
}

// `off` is one exclusive scan over the grid cells of every index of the
// batch, so positions are global: index r holds [r*n, (r+1)*n).  The scatter
// writes one int per point; the coordinates and cells are then gathered in
// sorted order, so their writes are coalesced.
__global__ void k_b_scatter(DevData D, const BatchSub *__restrict__ subs, const int *__restrict__ reps,
                            int *__restrict__ keys, const int *__restrict__ rank, const int *__restrict__ off,
                            int *__restrict__ order) {
    const BatchSub &B = subs[reps[blockIdx.y]];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.n) return;
    const int at = off[B.hoff + keys[B.obase + p]] + rank[B.obase + p];
    order[at] = p;
    keys[B.obase + p] = at - (int) B.obase;  // from here on: the point's local sorted position
}

// the index builder's columns: every dimension but the child's own (all of
// them when nothing is shared), and the packed cells of the checked ones
// Sorted columns kept across batches: ysrc[c] (c = rep position * kmax + t
// for the index columns, the subspace for its own column) is a kept copy of
// the column, read in order instead of gathered; ystore[c] receives a newly
// gathered column to keep.  Either may be null.
__global__ void k_b_gather(DevData D, const BatchSub *__restrict__ subs, const int *__restrict__ reps,
                           const int *__restrict__ order, float *__restrict__ Ys, uint16_t *__restrict__ Es,
                           uint64_t *__restrict__ sigS, int need_es, const float *const *__restrict__ ysrc,
                           float *const *__restrict__ ystore, int kmax) {
    const BatchSub &B = subs[reps[blockIdx.y]];
    const int n = D.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = order[B.obase + i];
    const float ecs = ecell_width(D.ecells);
    uint64_t w = 0;
    for (int t = 0; t < B.S.k; ++t) {
        if (t == B.pos_m) continue;
        const int s = B.S.dims[t];
        const float *src = ysrc[blockIdx.y * kmax + t];
        float v;
        if (src) {
            v = src[i];
        } else {
            v = D.Xt[(size_t) s * n + p];
            float *dst = ystore[blockIdx.y * kmax + t];
            if (dst) dst[i] = v;
        }
        Ys[B.coord + (long long) t * n + i] = v;
        // the cells from the coordinate itself (k_discretise's formula on the
        // same value): no scattered read of the cell array
        int f = -1;
        for (int u = 0; u < B.nchk4; ++u)
            if (B.chk[u] == t) f = u;
        if (need_es || f >= 0) {
            const int e = ecell_of(v, ecs, D.ecells);
            if (need_es) Es[B.coord + (long long) t * n + i] = (uint16_t) e;
            if (f >= 0) w |= (uint64_t) e << (16 * f);
        }
    }
    sigS[B.obase + i] = w;
}

// every subspace: its own column and its packed cells
__global__ void k_b_gather_own(DevData D, const BatchSub *__restrict__ subs, const int *__restrict__ order,
                               float *__restrict__ Ys, uint16_t *__restrict__ Es, const uint64_t *__restrict__ sigS,
                               uint64_t *__restrict__ sig, int need_es, const float *const *__restrict__ ysrc,
                               float *const *__restrict__ ystore) {
    const BatchSub &B = subs[blockIdx.y];
    const int n = D.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t w = sigS[B.obase + i];
    if (B.pos_m < B.S.k) {
        const int p = order[B.obase + i];
        const int s = B.S.dims[B.pos_m];
        const float *src = ysrc[blockIdx.y];
        float v;
        if (src) {
            v = src[i];
        } else {
            v = D.Xt[(size_t) s * n + p];
            if (float *dst = ystore[blockIdx.y]) dst[i] = v;
        }
        Ys[B.mcoord + i] = v;
        bool own_checked = false;
        for (int f = 0; f < B.nchk4; ++f) own_checked = own_checked || B.chk[f] == B.pos_m;
        if (need_es || own_checked) {
            const uint16_t e = (uint16_t) ecell_of(v, ecell_width(D.ecells), D.ecells);
            if (need_es) Es[B.mcoord + i] = e;
            for (int f = 0; f < B.nchk4; ++f)
                if (B.chk[f] == B.pos_m) w |= (uint64_t) e << (16 * f);
        }
    }
    sig[B.base + i] = w;
}

// ------------------------------------------------------------ query lists in sorted order
// pos: every point's local sorted position in its index
__global__ void k_b_qmark(const BatchSub *__restrict__ subs, const QuerySeg *__restrict__ segs,
                          const int *__restrict__ blk_off, int nseg, const int *__restrict__ pos,
                          uint8_t *__restrict__ qflag) {
    const int s = seg_of(blk_off, nseg, blockIdx.x);
    const QuerySeg Q = segs[s];
    const int i = (blockIdx.x - blk_off[s]) * blockDim.x + threadIdx.x;
    if (i >= Q.len) return;
    const BatchSub &B = subs[Q.b];
    const int p = Q.P[i];
    qflag[B.base + pos[B.obase + p]] = 1;
}

// kept sorts: copied in after the scan of the indexes still to sort; sorts
// to keep are copied out of the batch after their scatter
__global__ void k_b_sort_in(const SortCopy *__restrict__ cs, int n, int *__restrict__ order, int *__restrict__ pos,
                            int *__restrict__ off) {
    const SortCopy &c = cs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        order[c.obase + i] = c.order[i];
        pos[c.obase + i] = c.pos[i];
    }
    if (i <= c.ncell) off[c.hoff + i] = c.off[i] + (int) c.obase;
}

__global__ void k_b_sort_out(const SortCopy *__restrict__ cs, int n, const int *__restrict__ order,
                             const int *__restrict__ pos, const int *__restrict__ off) {
    const SortCopy &c = cs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        c.order[i] = order[c.obase + i];
        c.pos[i] = pos[c.obase + i];
    }
    if (i <= c.ncell) c.off[i] = off[c.hoff + i] - (int) c.obase;
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

// ------------------------------------------------------------ candidate walk
// The candidates of a query are up to 9 runs of sorted positions (grid cells
// of the index dimensions).  They are concatenated into one index space
// [0, total): q[0..8] run starts (local positions), q[9..18] prefix sums of
// the run lengths, q[19] the number of runs.  A lane maps its candidate
// number to a position with a run pointer that only moves forward.
// q[20]: the run of the query's own cells in the first index dimensions (it
// holds the query, so it is not empty).  q[21..30]: prefix sums of the runs'
// lengths in groups of four (ceil(length / 4)), for the walks that take four
// consecutive candidates of one run per lane.
__device__ __forceinline__ void build_runs(const GridArg &G, const int *__restrict__ off, const int *qe,
                                           const float *qy, int stride, float s_star, int base0, int *q) {
    // rows r = 3 xi + yi over the cells within one of the query's in the first
    // (and second) index dimension, each a run of the last one's cells within
    // one; every row's two offsets are loaded before any is used, so the
    // set-up costs one memory round trip
    const int C = G.C;
    const int c0 = qe[G.tI[0] * stride];
    const int c1 = G.j == 3 ? qe[G.tI[1] * stride] : 0;
    const int cl = qe[G.tI[G.j - 1] * stride];
    const int F = G.F > 1 ? G.F : 1, CF = C * F;
    // a run: the buckets of the last dimension's cells within one of the
    // query's, less those farther than eps from the query's coordinate (a
    // neighbour's squared difference there is at most s_star; the bounds are
    // rounded outwards and the bucket map is monotone, so no neighbour is lost)
    int lo = max(cl - 1, 0) * F, hi = min(cl + 1, C - 1) * F + F - 1;
    if (F > 1) {
        const double xq = (double) qy[G.tI[G.j - 1] * stride];
        const double em = sqrt((double) s_star) * 1.001 + 1e-6;
        lo = max(lo, fine_key(__double2float_rd(xq - em), C, F));
        hi = min(hi, fine_key(__double2float_ru(xq + em), C, F));
    }
    const int x0 = max(c0 - 1, 0), nx = G.j == 1 ? 1 : min(c0 + 1, C - 1) - x0 + 1;
    const int y0 = max(c1 - 1, 0), ny = G.j == 3 ? min(c1 + 1, C - 1) - y0 + 1 : 1;
    const int own = G.j == 1 ? 0 : 3 * (c0 - x0) + (G.j == 3 ? c1 - y0 : 0);
    int a[9], e[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        // rows past the edge of the grid load a row inside it (no branch
        // around the loads) and are dropped below
        const int xi = min(r / 3, nx - 1), yi = min(r % 3, ny - 1);
        const int row = G.j == 1 ? 0 : G.j == 2 ? (x0 + xi) * CF : ((x0 + xi) * C + y0 + yi) * CF;
        a[r] = off[row + lo];
        e[r] = off[row + hi + 1];
    }
    int nr = 0, acc = 0, gacc = 0;
    q[9] = 0;
    q[20] = 0;
    q[21] = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        if (r / 3 < nx && r % 3 < ny && e[r] > a[r]) {
            if (r == own) q[20] = nr;
            q[nr] = a[r] - base0;
            acc += e[r] - a[r];
            gacc += (e[r] - a[r] + 3) >> 2;
            q[10 + nr] = acc;
            q[22 + nr] = gacc;
            ++nr;
        }
    }
    q[19] = nr;
}

__device__ __forceinline__ int run_pos(const int *q, int c, int &r) {
    while (c >= q[10 + r]) ++r;
    return q[r] + (c - q[9 + r]);
}

constexpr int kFastDims = 8;

// the packed cells of the checked dimensions lie within one cell of the query's
__device__ __forceinline__ bool cells_near(uint64_t w, uint64_t qsig, int nc) {
    bool ok = true;
    for (int f = 0; f < nc; ++f) {
        const int dc = (int) ((w >> (16 * f)) & 0xFFFFull) - (int) ((qsig >> (16 * f)) & 0xFFFFull);
        ok = ok && dc >= -1 && dc <= 1;
    }
    return ok;
}

// Candidate at local sorted position jj against the staged query: cells of the
// checked dimensions within one (packed word first), then the distance summed
// in subspace order with separate multiplies and adds, compared with s_star.
__device__ __forceinline__ bool cand_test(const BatchSub &B, int n, const float *__restrict__ Ys,
                                          const uint16_t *__restrict__ Es, const uint64_t *__restrict__ sigb,
                                          uint64_t qsig, const float *qy, const int *qe, int stride, int jj,
                                          float s_star, float &acc_out) {
    const int nc = B.nchk4;
    if (nc > 0 && !cells_near(sigb[jj], qsig, nc)) return false;
    const uint64_t rest = B.restmask;
    const int k = B.S.k;
    float acc = 0.0f;
    for (int t = 0; t < k; ++t) {
        const long long col = col_of(B, t, n);
        if ((rest >> t) & 1ull) {
            const int de = (int) Es[col + jj] - qe[t * stride];
            if (de < -1 || de > 1) return false;
        }
        const float diff = qy[t * stride] - Ys[col + jj];
        const float sq = diff * diff;
        acc = acc + sq;
        if (acc > s_star) return false;
    }
    acc_out = acc;
    return true;
}

// LDS of the neighbour passes: per query k coordinates, k cells, its packed
// cells (2 words) and its runs
__device__ __forceinline__ void lds_layout(char *smem, int k, float *&sy, int *&se, uint64_t *&qs, int *&qr) {
    sy = reinterpret_cast<float *>(smem);
    se = reinterpret_cast<int *>(smem + sizeof(float) * QPB * k);
    qs = reinterpret_cast<uint64_t *>(smem + 8 * (size_t) QPB * k);
    qr = reinterpret_cast<int *>(smem + 8 * (size_t) QPB * k + 8 * QPB);
}

// the query's coordinates (sorted position i) and cells (point p) into LDS:
// the first 8 dimensions' loads are all issued before any store (a store to
// LDS through a generic pointer would otherwise order each load after the
// previous store, one memory round trip per dimension)
// (the cells are computed from the coordinates with the discretisation's
// formula, G.C being the number of eps cells: no scattered read by point id)
__device__ __forceinline__ void stage_query_b(const BatchSub &B, int n, const float *__restrict__ Ys, int p, int i,
                                              float *sy, int *se, int stride) {
    constexpr int kStage = 8;
    const int k = B.S.k;
    const float ecs = ecell_width(B.G.C);
    (void) p;
    float y[kStage];
#pragma unroll
    for (int t = 0; t < kStage; ++t) y[t] = t < k ? Ys[col_of(B, t, n) + i] : 0.0f;
#pragma unroll
    for (int t = 0; t < kStage; ++t)
        if (t < k) {
            sy[t * stride] = y[t];
            se[t * stride] = ecell_of(y[t], ecs, B.G.C);
        }
    for (int t = kStage; t < k; ++t) {
        const float v = Ys[col_of(B, t, n) + i];
        sy[t * stride] = v;
        se[t * stride] = ecell_of(v, ecs, B.G.C);
    }
}

// ------------------------------------------------------------ candidate walks of one query
// The LPQ lanes of a query take candidates c0 + u*LPQ + sub of each step.
// The walk starts at a pseudo-random candidate of the run of the query's own
// cells (the densest part of its neighbourhood) and wraps: the first
// neighbours found are then spread over the candidates (random sample edges
// for the components pass, as Afforest draws them) instead of the same few
// points at the start of a shared run.
struct QueryWalk {
    const BatchSub &B;
    int n, i, p;
    const int *q;    // the query's runs (LDS)
    const float *qy; // its coordinates (LDS, stride QPB)
    const int *qe;   // its cells (LDS, stride QPB)
    uint64_t qsig;   // its packed cells
    const int *__restrict__ order;
    const float *__restrict__ Ys;
    const uint16_t *__restrict__ Es;
    const uint64_t *__restrict__ sigb;
    float s_star, eps;
    double phirel;
    int *nbp;        // its two recorded neighbours
};

__device__ __forceinline__ int walk_rot(const QueryWalk &W, int total) {
    const int *q = W.q;
    const int cs0 = q[9 + q[20]], clen = q[10 + q[20]] - cs0;
    return total > 0 ? cs0 + (int) (((uint32_t) W.p * 2654435761u) % (uint32_t) max(clen, 1)) : 0;
}

// per 16-bit field (w - q + 1) mod 2^16 is 0, 1 or 2 exactly when the cells
// differ by at most one (cells < 65535); fields beyond the checked
// dimensions are zero in both words and pass
__device__ __forceinline__ bool cells_near_packed(uint64_t w, uint64_t qsig) {
    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    const u16x2 one = {1, 1};
    const u16x2 d0 = __builtin_bit_cast(u16x2, (uint32_t) w) - __builtin_bit_cast(u16x2, (uint32_t) qsig) + one;
    const u16x2 d1 =
        __builtin_bit_cast(u16x2, (uint32_t) (w >> 32)) - __builtin_bit_cast(u16x2, (uint32_t) (qsig >> 32)) + one;
    const u16x2 m = __builtin_elementwise_max(d0, d1);
    return m.x <= 2 && m.y <= 2;
}

// neighbours found in one step: the first kSampleEdges are recorded (the sample edges
// of the components pass); phi adds the terms of the step
template <int U, bool PHI>
__device__ __forceinline__ void walk_tally(const QueryWalk &W, const int (&jj)[U], const bool (&hit)[U],
                                           const float (&accs)[U], int &found, double &dsum) {
    constexpr int LPQ = kBLanes;
    const int sub = __lane_id() % LPQ;
    const int gshift = (__lane_id() / LPQ) * LPQ;
    const unsigned long long below = (1ull << sub) - 1ull;
    if (PHI) {
        double part = 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (hit[u]) part += phi_term(accs[u], W.eps);
        for (int o = LPQ / 2; o > 0; o >>= 1) part += __shfl_xor(part, o);
        dsum += part;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const unsigned long long mine = (__ballot(hit[u]) >> gshift) & ((1ull << LPQ) - 1ull);
        if (hit[u] && found < kSampleEdges) {
            const int at = found + __popcll(mine & below);
            if (at < kSampleEdges) W.nbp[at] = W.order[W.B.obase + jj[u]];
        }
        found += __popcll(mine);
    }
}

__device__ __forceinline__ bool walk_more(const QueryWalk &W, bool phi, int found, double dsum, int total) {
    if (phi) {
        // phi may stop once its running double sum is certainly above both thresholds
        const double thr_max = (double) fmaxf(W.B.phi_w, W.B.phi_d);
        return dsum * (1.0 - (double) total * W.phirel) < thr_max;
    }
    return found < W.B.cap;
}

// k = K dimensions, every checked one packed (the host checks), counts or
// phi.  Each lane takes four consecutive candidates of one run per step, so a
// step loads each coordinate column with one 16-byte load per lane (and the
// packed cells with two) instead of four 4-byte loads.  The vector memory pipeline bounds
// these walks (its address and data units are busy nearly all the time),
// and its cost grows with the number of load instructions.  The runs are
// concatenated in groups of four (q[21..30]); the last group of a run masks
// the positions past the run's end, which are read (up to three past the
// run: the batch's coordinate and packed-cell buffers end with padding) but
// never counted.  A lane's first group position is 4-byte aligned only;
// dword-aligned multi-dword global loads are supported on gfx950.
template <int K, bool PHI>
__device__ __forceinline__ void walk_vec(const QueryWalk &W, int &found, double &dsum, bool &more, long long &seen) {
    typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
    typedef unsigned long long u2u __attribute__((ext_vector_type(2), aligned(8)));
    constexpr int LPQ = kBLanes, U = 4;
    const BatchSub &B = W.B;
    const int *q = W.q;
    const int sub = __lane_id() % LPQ;
    const int nr = q[19];
    const int gtotal = q[21 + nr], total = q[9 + nr];
    // start at a pseudo-random group of the query's own run (spreads the
    // sample edges), walk to the end, then from the start to there
    const int own = q[20];
    const int og = q[21 + own], olen = q[22 + own] - og;
    const int grot = gtotal > 0 ? og + (int) (((uint32_t) W.p * 2654435761u) % (uint32_t) max(olen, 1)) : 0;
    const float *col[K];
    float qv[K];
#pragma unroll
    for (int t = 0; t < K; ++t) {
        col[t] = W.Ys + (t == B.pos_m ? B.mcoord : B.coord + (long long) t * W.n);
        qv[t] = W.qy[t * QPB];
    }
    for (int segm = 0; segm < 2; ++segm) {
        const int gbeg = segm == 0 ? grot : 0, gend = segm == 0 ? gtotal : grot;
        int r = segm == 0 ? own : 0;
        int rgend = q[22 + r];
        for (int g0 = gbeg; g0 < gend && more; g0 += LPQ) {
            const int g = min(g0 + sub, gend - 1);
            while (g >= rgend) {
                ++r;
                rgend = q[22 + r];
            }
            const int e0 = (g - q[21 + r]) << 2;  // the group's first element in run r
            const int len = q[10 + r] - q[9 + r];
            const int nv = g0 + sub < gend ? min(U, len - e0) : 0;  // its elements inside the run
            const int jj0 = q[r] + e0;
            const u2u wa = *reinterpret_cast<const u2u *>(W.sigb + (uint32_t) jj0);
            const u2u wb = *reinterpret_cast<const u2u *>(W.sigb + (uint32_t) jj0 + 2);
            f4u y[K];
#pragma unroll
            for (int t = 0; t < K; ++t)
                y[t] = *reinterpret_cast<const f4u *>(reinterpret_cast<const char *>(col[t]) +
                                                      ((uint32_t) jj0 << 2));
            const uint64_t w[U] = {wa.x, wa.y, wb.x, wb.y};
            int jj[U];
            bool hit[U];
            float accs[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < K; ++t) {
                    const float diff = qv[t] - y[t][u];
                    const float sq = diff * diff;
                    acc = acc + sq;
                }
                jj[u] = jj0 + u;
                hit[u] = (u < nv) & (jj[u] != W.i) & cells_near_packed(w[u], W.qsig) & (acc <= W.s_star);
                accs[u] = acc;
            }
            // the query's examined candidates of the step, summed over its
            // quad with two DPP moves
            int ex = nv;
            ex += __builtin_amdgcn_update_dpp(0, ex, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
            ex += __builtin_amdgcn_update_dpp(0, ex, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
            // the query's hits, summed over the quad with two DPP moves (phi:
            // and the step's phi terms); the ballots of walk_tally only while
            // the two sample edges are open
            int h = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) h += hit[u] ? 1 : 0;
            h += __builtin_amdgcn_update_dpp(0, h, 0xB1, 0xF, 0xF, false);
            h += __builtin_amdgcn_update_dpp(0, h, 0x4E, 0xF, 0xF, false);
            if (PHI) {
                double part = 0.0;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (hit[u]) part += phi_term(accs[u], W.eps);
                part += __shfl_xor(part, 1);
                part += __shfl_xor(part, 2);
                dsum += part;
            }
            if (found < kSampleEdges && h > 0) {
                double unused = 0.0;
                int f = found;
                walk_tally<U, false>(W, jj, hit, accs, f, unused);
            }
            found += h;
            seen += ex;
            more = walk_more(W, PHI, found, dsum, total);
        }
    }
}

// k <= kFastDims with every checked dimension packed, k known only at run
// time (batches that mix dimensionalities): loads predicated on the packed
// cell test, the sum over the first k of kFastDims terms
template <bool PHI>
__device__ __forceinline__ void walk_fast_rt(const QueryWalk &W, int &found, double &dsum, bool &more,
                                             long long &seen) {
    constexpr int LPQ = kBLanes, U = 4, STEP = LPQ * U;
    const BatchSub &B = W.B;
    const int *q = W.q;
    const int sub = __lane_id() % LPQ;
    const int total = q[9 + q[19]];
    const int rot = walk_rot(W, total);
    const int k = B.S.k, nc = B.nchk4, pm = B.pos_m;
    const long long cb = B.coord, mc = B.mcoord;
    float qv[kFastDims];
#pragma unroll
    for (int t = 0; t < kFastDims; ++t) qv[t] = t < k ? W.qy[t * QPB] : 0.0f;
    int r = 0;
    bool wrapped = false;
    for (int c0 = 0; c0 < total && more; c0 += STEP) {
        int jj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + u * LPQ + sub;
            jj[u] = -1;
            if (c < total) {
                int cr = c + rot;
                if (cr >= total) {
                    cr -= total;
                    if (!wrapped) {
                        wrapped = true;
                        r = 0;
                    }
                }
                jj[u] = run_pos(q, cr, r);
            }
        }
        uint64_t w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = (jj[u] >= 0 && nc > 0) ? W.sigb[jj[u]] : 0ull;
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) ok[u] = jj[u] >= 0 && jj[u] != W.i && cells_near(w[u], W.qsig, nc);
        float y[U][kFastDims];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int t = 0; t < kFastDims; ++t)
                y[u][t] = (ok[u] && t < k) ? W.Ys[(t == pm ? mc : cb + (long long) t * W.n) + jj[u]] : 0.0f;
        bool hit[U];
        float accs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            float acc = 0.0f;
#pragma unroll
            for (int t = 0; t < kFastDims; ++t) {
                if (t < k) {
                    const float diff = qv[t] - y[u][t];
                    const float sq = diff * diff;
                    acc = acc + sq;
                }
            }
            hit[u] = ok[u] && acc <= W.s_star;
            accs[u] = acc;
        }
        walk_tally<U, PHI>(W, jj, hit, accs, found, dsum);
        seen += min(STEP, total - c0);
        more = walk_more(W, PHI, found, dsum, total);
    }
}

// any k, cells of the checked dimensions beyond the packed four compared one
// by one (cand_test)
template <bool PHI>
__device__ __forceinline__ void walk_generic(const QueryWalk &W, int &found, double &dsum, bool &more,
                                             long long &seen) {
    constexpr int LPQ = kBLanes, U = 4, STEP = LPQ * U;
    const int *q = W.q;
    const int sub = __lane_id() % LPQ;
    const int total = q[9 + q[19]];
    const int rot = walk_rot(W, total);
    int r = 0;
    bool wrapped = false;
    for (int c0 = 0; c0 < total && more; c0 += STEP) {
        int jj[U];
        bool hit[U];
        float accs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + u * LPQ + sub;
            jj[u] = -1;
            if (c < total) {
                int cr = c + rot;
                if (cr >= total) {
                    cr -= total;
                    if (!wrapped) {
                        wrapped = true;
                        r = 0;
                    }
                }
                jj[u] = run_pos(q, cr, r);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            accs[u] = 0.0f;
            hit[u] = jj[u] >= 0 && jj[u] != W.i &&
                     cand_test(W.B, W.n, W.Ys, W.Es, W.sigb, W.qsig, W.qy, W.qe, QPB, jj[u], W.s_star, accs[u]);
        }
        walk_tally<U, PHI>(W, jj, hit, accs, found, dsum);
        seen += min(STEP, total - c0);
        more = walk_more(W, PHI, found, dsum, total);
    }
}

// ------------------------------------------------------------ neighbour counts
// k_grid_count over a batch: LIST = the compacted queries of each subspace
// (blocks by segment), else every sorted position of subspace blockIdx.y whose
// count is still missing (components pass, with the subset bound).
template <bool LIST, bool PHI, int K>
__global__ __launch_bounds__(kBlock) void k_b_count(const BatchSub *__restrict__ subs, const int *__restrict__ blk_off,
                                                    int nsub, int n, const int *__restrict__ off,
                                                    const int *__restrict__ order, const float *__restrict__ Ys,
                                                    const uint16_t *__restrict__ Es, const uint64_t *__restrict__ sig,
                                                    const int *__restrict__ qlist, const int *__restrict__ qstart,
                                                    float s_star, float eps, double phirel, int *__restrict__ nb,
                                                    int2 *__restrict__ fb, int *__restrict__ nfb, int fbcap,
                                                    unsigned long long *__restrict__ stat) {
    constexpr int LPQ = kBLanes;
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
    float *sy;
    int *se, *qr;
    uint64_t *qs;
    lds_layout(smem, k, sy, se, qs, qr);
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
    const uint64_t *sigb = sig + B.base;
    int *cnt = B.cnt;
    const int p = active ? order[B.obase + i] : 0;
    active = active && cnt[p] < 0;
    // N_S(p) is a subset of N_S'(p) for S' a subset of S: the smallest
    // complete count of the subspaces S minus one dimension, when below both
    // thresholds of S, decides p without a scan (and stays an upper bound);
    // phi: the smallest stored upper bound, below both phi thresholds
    if (active && B.nbound > 0) {
        int bb = INT_MAX;
        for (int t = 0; t < B.nbound; ++t) {
            const int v = B.bound[t][p];
            if (v >= 0) bb = min(bb, PHI ? v >> 2 : v);
        }
        if (PHI ? bb < kPhiNoBound && (double) bb * 0.25 < (double) fminf(B.phi_w, B.phi_d) : bb < B.lo) {
            if (sub == 0) cnt[p] = PHI ? bb << 2 : bb;
            active = false;
        }
    }
    const int base0 = (int) B.obase;
    int *q = qr + lq * kRangeWords;
    if (sub == 0 && active) {
        const uint64_t qsg = sigb[i];
        stage_query_b(B, n, Ys, p, i, sy + lq, se + lq, QPB);
        qs[lq] = qsg;
        build_runs(B.G, off + B.hoff, se + lq, sy + lq, QPB, s_star, base0, q);
    }
    __syncthreads();
    long long seen = 0;
    int found = 0;
    if (active) {
        double dsum = 0.0;
        bool more = true;
        const QueryWalk W{B, n, i, p, q, sy + lq, se + lq, qs[lq], order, Ys, Es, sigb, s_star, eps, phirel,
                          nb + kSampleEdges * (B.base + p)};
        // K > 0: every subspace of the launch has k = K with its checked
        // dimensions packed (the host checks), so the walk is compiled for it
        if constexpr (K > 0)
            walk_vec<K, PHI>(W, found, dsum, more, seen);
        else if (k <= kFastDims && B.restmask == 0)
            walk_fast_rt<PHI>(W, found, dsum, more, seen);
        else
            walk_generic<PHI>(W, found, dsum, more, seen);
        if (PHI) {
            // decision bits: 1 weak-dense, 2 dense; a threshold inside the
            // rounding band sends the point to the leaf-order recomputation
            const double lo = dsum * (1.0 - (double) found * phirel);
            const double hi = dsum * (1.0 + (double) found * phirel);
            int code = 0;
            bool unsure = false;
            double ub = -1.0;  // stopped early: no bound
            if (!more) {
                code = 3;
            } else {
                const double tw = (double) B.phi_w, td = (double) B.phi_d;
                if (lo >= tw) code |= 1;
                else if (!(hi < tw)) unsure = true;
                if (lo >= td) code |= 2;
                else if (!(hi < td)) unsure = true;
                // hi bounds T, the exact sum of the terms (the double scan-order
                // sum is within N 2^-52 of it)
                ub = hi * (1.0 + (double) found * phirel);
            }
            if (sub == 0) {
                cnt[p] = phi_code(code, ub);
                if (unsure) {
                    const int at = atomicAdd(nfb, 1);
                    if (at < fbcap) fb[at] = make_int2(b, p);
                }
            }
        } else if (sub == 0) {
            cnt[p] = found;
        }
    }
    if (stat) {
        // per wave into one of 64 slot groups: [0] candidates examined, [1]
        // points counted, [2] algorithmic bytes (4k + 4 per point, 4 + 4k per
        // neighbour established), [3] bytes of the examined candidates (their
        // packed cells when packed, and k coordinates)
        const bool mine = sub == 0 && active;
        long long done = mine ? 1 : 0;
        long long abytes = mine ? (4ll * k + 4) + (long long) found * (4 + 4ll * k) : 0;
        if (sub != 0) seen = 0;
        long long ebytes = seen * ((B.nchk4 > 0 ? 8ll : 0ll) + 4ll * k);
        for (int o = 32; o > 0; o >>= 1) {
            seen += __shfl_down(seen, o);
            done += __shfl_down(done, o);
            abytes += __shfl_down(abytes, o);
            ebytes += __shfl_down(ebytes, o);
        }
        if (__lane_id() == 0 && done) {
            unsigned long long *slot = stat + 16 * (blockIdx.x & 63);
            atomicAdd(&slot[0], (unsigned long long) seen);
            atomicAdd(&slot[1], (unsigned long long) done);
            atomicAdd(&slot[2], (unsigned long long) abytes);
            atomicAdd(&slot[3], (unsigned long long) ebytes);
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
    // i is a sorted position for the dense flag and a point id for the
    // parents and sizes, so every write is coalesced
    const int p = order[B.obase + i];
    dsorted[B.base + i] = meets(B.cnt[p], B.thr_dense, B.dbit) ? 1 : 0;
    par[B.base + i] = meets(B.cnt[i], B.thr_dense, B.dbit) ? i : -1;
    B.csize[i] = 0;
}

// Sample phase of the components: every dense point links to the dense ones
// among the first kSampleEdges neighbours its count pass recorded (any edges would do:
// the full phase below links every point outside the giant component to all
// its dense neighbours, so the result does not depend on the sample).
// One round per recorded neighbour, with the roots compressed in between, so
// that the finds of the second round walk short paths (Afforest's rounds).
__global__ void k_b_link_sample(const BatchSub *__restrict__ subs, int n, const int *__restrict__ nb, int round,
                                int *__restrict__ par) {
    const BatchSub &B = subs[blockIdx.y];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    // a point is dense exactly when its parent is set (k_b_dense_init gives
    // dense points themselves, the others -1, and unions only move roots to
    // smaller roots), so the parents answer both density tests without the
    // counts
    int *pr = par + B.base;
    const int pp = pr[p];
    if (pp < 0) return;
    // the recorded neighbours: min(count, kSampleEdges) of them, -1 after
    // (phi mode records them too)
    const int q = nb[kSampleEdges * (B.base + p) + round];
    if (q < 0 || q >= n) return;
    // after the first round's compression every point's parent is its root:
    // equal parents are one set, decided without reading the root's own entry
    // (which every point of a large component would read at once)
    const int pq = pr[q];
    if (pq < 0 || pq == pp) return;
    uf_union(pr, p, q);
}

__global__ void k_b_uf_compress(const BatchSub *__restrict__ subs, int n, int *__restrict__ par) {
    int *pr = par + subs[blockIdx.y].base;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && pr[i] >= 0) pr[i] = uf_find(pr, i);
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
    roots[t] = p < n ? pr[p] : -1;
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

// Full phase: every dense point outside the giant component links to all its
// dense neighbours.  K > 0: every subspace of the launch has k = K with its
// checked dimensions packed, and a step loads the dense flags, packed cells,
// point ids and K coordinates of U candidates per lane at once (a lane past
// the end re-reads the last candidate and is masked).
template <int K>
__global__ __launch_bounds__(kBlock) void k_b_link_full(const BatchSub *__restrict__ subs, int n,
                                                        const int *__restrict__ off, const int *__restrict__ order,
                                                        const float *__restrict__ Ys, const uint16_t *__restrict__ Es,
                                                        const uint64_t *__restrict__ sig,
                                                        const uint8_t *__restrict__ dsorted, float s_star,
                                                        const int *__restrict__ giant, int *__restrict__ par) {
    constexpr int LPQ = kBLanes;
    extern __shared__ char smem[];
    const BatchSub &B = subs[blockIdx.y];
    const int k = B.S.k;
    float *sy;
    int *se, *qr;
    uint64_t *qs;
    lds_layout(smem, k, sy, se, qs, qr);
    const int lq = threadIdx.x / LPQ, sub = threadIdx.x % LPQ;
    const int i = blockIdx.x * QPB + lq;
    const uint8_t *ds = dsorted + B.base;
    const int *ord = order + B.obase;
    const uint64_t *sigb = sig + B.base;
    int *pr = par + B.base;
    bool active = i < n && ds[i];
    const int p = active ? ord[i] : 0;
    if (active && pr[p] == giant[blockIdx.y]) active = false;
    int *q = qr + lq * kRangeWords;
    if (sub == 0 && active) {
        const uint64_t qsg = sigb[i];
        stage_query_b(B, n, Ys, p, i, sy + lq, se + lq, QPB);
        qs[lq] = qsg;
        build_runs(B.G, off + B.hoff, se + lq, sy + lq, QPB, s_star, (int) B.obase, q);
    }
    __syncthreads();
    if (!active) return;
    const float *qy = sy + lq;
    const int *qe = se + lq;
    const uint64_t qsig = qs[lq];
    const int total = q[9 + q[19]];
    int rp = uf_find(pr, p);
    if constexpr (K > 0) {
        // each lane four consecutive candidates of one run per step (the
        // groups of walk_vec): per step one 16-byte load for the ids and each
        // coordinate column, two for the packed cells and two aligned words
        // for the four dense flags
        typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
        typedef int i4u __attribute__((ext_vector_type(4), aligned(4)));
        typedef unsigned long long u2u __attribute__((ext_vector_type(2), aligned(8)));
        constexpr int U = 4;
        const float *col[K];
        float qv[K];
#pragma unroll
        for (int t = 0; t < K; ++t) {
            col[t] = Ys + (t == B.pos_m ? B.mcoord : B.coord + (long long) t * n);
            qv[t] = qy[t * QPB];
        }
        const uint32_t *dw = reinterpret_cast<const uint32_t *>(dsorted);  // 256-byte aligned
        const int nr = q[19];
        const int gtotal = q[21 + nr];
        int r = 0, rgend = q[22];
        for (int g0 = 0; g0 < gtotal; g0 += LPQ) {
            const int g = min(g0 + sub, gtotal - 1);
            while (g >= rgend) {
                ++r;
                rgend = q[22 + r];
            }
            const int e0 = (g - q[21 + r]) << 2;
            const int len = q[10 + r] - q[9 + r];
            const int nv = g0 + sub < gtotal ? min(U, len - e0) : 0;
            const int jj0 = q[r] + e0;
            const size_t at = (size_t) B.base + (size_t) jj0;  // byte index of the first dense flag
            const uint32_t d0 = dw[at >> 2], d1 = dw[(at >> 2) + 1];
            const u2u wa = *reinterpret_cast<const u2u *>(sigb + (uint32_t) jj0);
            const u2u wb = *reinterpret_cast<const u2u *>(sigb + (uint32_t) jj0 + 2);
            const i4u qv4 = *reinterpret_cast<const i4u *>(ord + (uint32_t) jj0);
            f4u y[K];
#pragma unroll
            for (int t = 0; t < K; ++t) y[t] = *reinterpret_cast<const f4u *>(col[t] + (uint32_t) jj0);
            const uint32_t dn4 = __builtin_amdgcn_alignbyte(d1, d0, (uint32_t) (at & 3));
            const uint64_t w[U] = {wa.x, wa.y, wb.x, wb.y};
            const int qq[U] = {qv4.x, qv4.y, qv4.z, qv4.w};
            bool hit[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < K; ++t) {
                    const float diff = qv[t] - y[t][u];
                    const float sq = diff * diff;
                    acc = acc + sq;
                }
                hit[u] = (u < nv) & (jj0 + u != i) & (((dn4 >> (8 * u)) & 0xFFu) != 0) &
                         cells_near_packed(w[u], qsig) & (acc <= s_star);
            }
            // parents only move up to roots: a parent equal to p's root is p's set
            int pq[U];
#pragma unroll
            for (int u = 0; u < U; ++u) pq[u] = hit[u] ? pr[qq[u]] : rp;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!hit[u] || pq[u] == rp || uf_find(pr, qq[u]) == rp) continue;
                uf_union(pr, p, qq[u]);
                rp = uf_find(pr, p);
            }
        }
    } else {
        int r = 0;
        for (int c = sub; c < total; c += LPQ) {
            const int jj = run_pos(q, c, r);
            float acc;
            if (jj == i || !ds[jj] || !cand_test(B, n, Ys, Es, sigb, qsig, qy, qe, QPB, jj, s_star, acc)) continue;
            const int qq = ord[jj];
            // parents only move up to roots: a parent equal to p's root is p's set
            if (pr[qq] == rp || uf_find(pr, qq) == rp) continue;
            uf_union(pr, p, qq);
            rp = uf_find(pr, p);
        }
    }
}

__global__ void k_b_uf_flatten(const BatchSub *__restrict__ subs, int n, int *__restrict__ par) {
    const BatchSub &B = subs[blockIdx.y];
    int *pr = par + B.base;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1;
    if (i < n && pr[i] >= 0) c = uf_find(pr, i);
    wave_add_by_key(B.csize, c, c >= 0);
    if (i < n) B.comp[i] = c;
}

// ------------------------------------------------------------ partition of every (tree, dimension) pair
// k_part_count / k_part_scatter per pair.  Only the weak-dense points and the
// paths are written; a path's origin keeps its "point of this tree" bit when
// the origin survives the weak-density prune of the group.
__global__ void k_b_part_count(const PartPair *__restrict__ pairs, const int *__restrict__ blk_off, int np,
                               const uint8_t *__restrict__ cellC, int n, int thr, int wbit) {
    __shared__ int h[3 * 256];
    const int j = seg_of(blk_off, np, blockIdx.x);
    const PartPair &A = pairs[j];
    const int blk = blockIdx.x - blk_off[j];
    const int G = A.G;
    const uint8_t *cm = cellC + (size_t) A.m * n;
    for (int t = threadIdx.x; t < 3 * G; t += blockDim.x) h[t] = 0;
    __syncthreads();
    if (blk < A.bp) {
        for (int u = 0; u < kPartPerThread; ++u) {
            const int i = (blk * kPartPerThread + u) * blockDim.x + threadIdx.x;
            if (i < A.nP) {
                const int p = A.P[i];
                const int g = A.gid[cm[p]];
                atomicAdd(&h[g], 1);
                if (meets(A.weakcnt[p], thr, wbit)) atomicAdd(&h[G + g], 1);
            }
        }
    } else {
        for (int u = 0; u < kPartPerThread; ++u) {
            const int i = ((blk - A.bp) * kPartPerThread + u) * blockDim.x + threadIdx.x;
            if (i < A.nsp) {
                const dmask_t mk = mload(A.spm, i, A.mw) & dabove(A.m);
                if (mk) atomicAdd(&h[2 * G + A.gid[cm[origin(A.spx[i])]]], 1);
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * G; t += blockDim.x)
        if (h[t]) atomicAdd(&A.cnt[t], h[t]);
}

// Only the groups whose weak-dense points reach min_size become surviving
// trees; the lists of the others are never read, so they are not written (a
// pair without such a group does nothing).  Slots: LDS ranks within the block,
// then one global reservation per (block, group).  Each thread takes
// kPartPerThread elements, one block-width apart.
__global__ void k_b_part_scatter(const PartPair *__restrict__ pairs, const int *__restrict__ blk_off, int np,
                                 const uint8_t *__restrict__ cellC, int n, int thr, int wbit, int min_size) {
    __shared__ int base[3 * 256];
    __shared__ int lcnt[3 * 256];
    __shared__ int lres[3 * 256];
    __shared__ int any_pass;
    const int j = seg_of(blk_off, np, blockIdx.x);
    const PartPair &A = pairs[j];
    const int blk = blockIdx.x - blk_off[j];
    const int G = A.G;
    const uint8_t *cm = cellC + (size_t) A.m * n;
    if (threadIdx.x == 0) any_pass = 0;
    for (int t = threadIdx.x; t < 3 * G; t += blockDim.x) lcnt[t] = 0;
    __syncthreads();
    for (int g = threadIdx.x; g < G; g += blockDim.x)
        if (A.cnt[G + g] >= min_size) any_pass = 1;
    if (threadIdx.x < 3) {
        int acc = 0;
        for (int g = 0; g < G; ++g) {
            base[threadIdx.x * G + g] = acc;
            acc += A.cnt[threadIdx.x * G + g];
        }
    }
    __syncthreads();
    if (!any_pass) return;  // uniform over the block
    int *cur = A.cnt + 3 * G;
    const bool pts = blk < A.bp;
    int key[kPartPerThread], rnk[kPartPerThread], val[kPartPerThread];
    dmask_t mk[kPartPerThread];
    for (int u = 0; u < kPartPerThread; ++u) {
        key[u] = -1;
        rnk[u] = 0;
        val[u] = 0;
        mk[u] = 0;
        if (pts) {
            const int i = (blk * kPartPerThread + u) * blockDim.x + threadIdx.x;
            if (i < A.nP) {
                const int p = A.P[i];
                const int g = A.gid[cm[p]];
                if (A.cnt[G + g] >= min_size && meets(A.weakcnt[p], thr, wbit)) {
                    key[u] = G + g;
                    val[u] = p;
                }
            }
        } else {
            const int i = ((blk - A.bp) * kPartPerThread + u) * blockDim.x + threadIdx.x;
            if (i < A.nsp) {
                mk[u] = mload(A.spm, i, A.mw) & dabove(A.m);
                const int xr = A.spx[i];
                const int xo = origin(xr);
                const int g = A.gid[cm[xo]];
                if (mk[u] && A.cnt[G + g] >= min_size) {
                    key[u] = 2 * G + g;
                    val[u] = (xr >= 0 && meets(A.weakcnt[xo], thr, wbit)) ? xo : ~xo;
                }
            }
        }
        if (key[u] >= 0) rnk[u] = atomicAdd(&lcnt[key[u]], 1);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * G; t += blockDim.x)
        if (lcnt[t]) lres[t] = atomicAdd(&cur[t], lcnt[t]);
    __syncthreads();
    for (int u = 0; u < kPartPerThread; ++u) {
        if (key[u] < 0) continue;
        const int pos = base[key[u]] + lres[key[u]] + rnk[u];
        if (pts) {
            A.outW[pos] = val[u];
        } else {
            A.outx[pos] = val[u];
            mstore(A.outm, pos, mk[u], A.mw);
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
    dmask_t m = 0;
    if (i < P.nsp && P.spx[i] < 0) m = mload(P.spm, i, P.mw);
    wave_or_mask(m, P.bits);
}

__global__ void k_b_path_insert(DevData D, const PruneChild *__restrict__ C, const int *__restrict__ blk_w, int nc) {
    const int c = seg_of(blk_w, nc, blockIdx.x);
    const PruneChild &P = C[c];
    const int i = (blockIdx.x - blk_w[c]) * blockDim.x + threadIdx.x;
    if (i >= P.nW) return;
    const dmask_t need = *P.bits;
    if (!need) return;
    const int w = P.W[i];
    const dmask_t smask = P.smask;
    uint64_t h = 0x243F6A8885A308D3ull;
    for (int t = 0; t < D.d; ++t) {
        const int k = D.level[t];
        if (dhas(smask, k)) continue;
        h = mix64(h + D.cellC[(size_t) k * D.n + w] + 1);
        if (!dhas(need, k)) continue;
        const uint64_t e = mix64(h ^ (0x9E3779B97F4A7C15ull * (uint64_t) (k + 1)));
        const uint64_t entry = (e & 0xFFFFFFFF00000000ull) | (uint32_t) w;
        uint64_t slot = e & P.tmask;
        while (true) {
            // a slot is written once, by the swap from empty: a plain load that
            // still sees it empty is corrected by the failed swap, a full one is final
            uint64_t cur = P.table[slot];
            if (cur == kEmpty) {
                cur = atomicCAS((unsigned long long *) &P.table[slot], (unsigned long long) kEmpty,
                                (unsigned long long) entry);
                if (cur == kEmpty) break;
            }
            if ((cur >> 32) == (entry >> 32) && same_prefix(D, smask, w, (int) (uint32_t) cur, t)) break;
            slot = (slot + 1) & P.tmask;
        }
    }
}

__global__ void k_b_path_check(DevData D, const PruneChild *__restrict__ C, const int *__restrict__ blk_p, int nc) {
    const int c = seg_of(blk_p, nc, blockIdx.x);
    const PruneChild &P = C[c];
    const int i = (blockIdx.x - blk_p[c]) * blockDim.x + threadIdx.x;
    int x = 0;
    dmask_t m = 0;
    if (i < P.nsp) {
        x = P.spx[i];
        m = mload(P.spm, i, P.mw);
        if (x < 0) {
            const int xo = ~x;
            const dmask_t smask = P.smask;
            uint64_t h = 0x243F6A8885A308D3ull;
            dmask_t keep = 0;
            for (int t = 0; t < D.d; ++t) {
                const int k = D.level[t];
                if (dhas(smask, k)) continue;
                h = mix64(h + D.cellC[(size_t) k * D.n + xo] + 1);
                if (!dhas(m, k)) continue;
                const uint64_t e = mix64(h ^ (0x9E3779B97F4A7C15ull * (uint64_t) (k + 1)));
                uint64_t slot = e & P.tmask;
                while (true) {
                    const uint64_t cur = P.table[slot];  // complete: the inserts ran in an earlier launch
                    if (cur == kEmpty) break;
                    if ((cur >> 32) == (e >> 32) && same_prefix(D, smask, xo, (int) (uint32_t) cur, t)) {
                        keep |= dbit(k);
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
        mstore(P.outm, slot, m, P.mw);
    }
}

// ------------------------------------------------------------ phi in leaf order
// phi_S(p) exactly as the reference sums it: over p's neighbours in the order
// of the neighbourhood tree's leaves (rank_order: points sorted by their eps
// cells over all dimensions in tree level order, ties by id), float running
// sum with a double term (ClusteringCpu.cu:116-124).  One block per point:
// the block walks all points in that order, 256 at a time, finds the
// neighbours (candidate cells within one in every matched dimension of S,
// then the distance) and one thread adds their terms in order.
__global__ __launch_bounds__(kBlock) void k_b_phi_exact(DevData D, const BatchSub *__restrict__ subs,
                                                        const int2 *__restrict__ fb, const int *__restrict__ nfb,
                                                        int fbcap, const int *__restrict__ rank_order, float eps,
                                                        float s_star) {
    __shared__ double terms[kBlock];
    __shared__ int wcnt[kBlock / 64];
    __shared__ float qv[kMaxSubDims];
    __shared__ int qc[kMaxSubDims];
    const int count = min(*nfb, fbcap);
    const int n = D.n;
    const int lane = __lane_id(), wave = threadIdx.x / 64;
    for (int item = blockIdx.x; item < count; item += gridDim.x) {
        const int2 it = fb[item];
        const BatchSub &B = subs[it.x];
        const int p = it.y;
        const int k = B.S.k;
        if ((int) threadIdx.x < k) {
            const int s = B.S.dims[threadIdx.x];
            qv[threadIdx.x] = D.Xt[(size_t) s * n + p];
            qc[threadIdx.x] = D.cellE[(size_t) s * n + p];
        }
        __syncthreads();
        float sum = 0.0f;
        long long nterm = 0;
        for (int base = 0; base < n; base += kBlock) {
            const int rk = base + (int) threadIdx.x;
            bool hit = false;
            double term = 0.0;
            if (rk < n) {
                const int q = rank_order[rk];
                bool ok = q != p;
                for (int t = 0; t < k && ok; ++t) {
                    if (!((B.matchmask >> t) & 1ull)) continue;
                    const int de = (int) D.cellE[(size_t) B.S.dims[t] * n + q] - qc[t];
                    ok = de >= -1 && de <= 1;
                }
                if (ok) {
                    float acc = 0.0f;
                    for (int t = 0; t < k; ++t) {
                        const float diff = qv[t] - D.Xt[(size_t) B.S.dims[t] * n + q];
                        const float sq = diff * diff;
                        acc = acc + sq;
                    }
                    if (acc <= s_star) {
                        hit = true;
                        term = phi_term(acc, eps);
                    }
                }
            }
            const unsigned long long m = __ballot(hit);
            if (lane == 0) wcnt[wave] = __popcll(m);
            __syncthreads();
            int off0 = 0, tot = 0;
            for (int w = 0; w < kBlock / 64; ++w) {
                if (w < wave) off0 += wcnt[w];
                tot += wcnt[w];
            }
            if (hit) terms[off0 + __popcll(m & ((1ull << lane) - 1ull))] = term;
            __syncthreads();
            if (threadIdx.x == 0)
                for (int t = 0; t < tot; ++t) sum = (float) ((double) sum + terms[t]);
            nterm += tot;
            __syncthreads();
        }
        // the float sum bounds T within N 2^-22 (so T <= sum (1 + 2 N 2^-22)
        // while N 2^-22 <= 1/2), and the stored bound adds N 2^-22 once more
        const double rel = (double) nterm * (1.0 / 4194304.0);
        if (threadIdx.x == 0)
            B.cnt[p] = phi_code((sum >= B.phi_w ? 1 : 0) | (sum >= B.phi_d ? 2 : 0),
                                rel <= 0.25 ? (double) sum * (1.0 + 4.0 * rel) : -1.0);
        __syncthreads();
    }
}

// ------------------------------------------------------------ replay of pruneRedundancy and join
// The reference's depth-first order is replayed in stages: one launch per
// pass serves every step of a stage (the host groups the steps).  `list`
// holds the result-table rows whose subspace contains S (fixed once the
// expansion is done); a row counts while its nkept > 0, which the joins of
// earlier stages have set.  A tree is known by its index; tstamp[p] is the
// tree of S whose weak-dense points hold p.
//
// Both row walks (redundancy, join check) are planned in LDS: every block of
// a step reads a segment of up to kSegRows rows at once, cuts each row's walk
// into chunks of kChunk items and takes every nb-th chunk (nb: the step's
// blocks), so no block walks the rows one after another.
constexpr int kChunk = 2048;
constexpr int kSegRows = 1024;

struct RowPlan {
    int len[kSegRows];      // items of the row's walk (0: the row is not in the result)
    int pre[kSegRows + 1];  // chunks before each row
    uint8_t by_mem[kSegRows];
};

// loads rows [seg, seg + ns) of the list and their chunk prefix; returns the
// number of chunks.  by_mem: walk the row's members (else the tree's points,
// when nW is shorter; nW < 0 walks members always)
__device__ int plan_rows(const Entry *__restrict__ E, const int *__restrict__ list, int seg, int ns, int nW,
                         RowPlan &R, int *wtmp) {
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
        const Entry e = E[list[seg + i]];
        const bool bm = nW < 0 || e.nmem <= nW;
        R.len[i] = e.nkept > 0 ? (bm ? e.nmem : nW) : 0;
        R.by_mem[i] = bm ? 1 : 0;
    }
    __syncthreads();
    // exclusive prefix of chunk counts: 4 rows per thread, then a block scan
    constexpr int PER = kSegRows / kBlock;
    int c[PER], sum = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int i = threadIdx.x * PER + u;
        c[u] = i < ns ? (R.len[i] + kChunk - 1) / kChunk : 0;
        sum += c[u];
    }
    int incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (__lane_id() >= o) incl += v;
    }
    if (__lane_id() == 63) wtmp[threadIdx.x / 64] = incl;
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < (int) threadIdx.x / 64) before += wtmp[w];
        total += wtmp[w];
    }
    int at = before + incl - sum;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int i = threadIdx.x * PER + u;
        if (i < ns) R.pre[i] = at;
        at += c[u];
    }
    if (threadIdx.x == 0) R.pre[ns] = total;
    __syncthreads();
    return total;
}

// row of chunk w: the last i with pre[i] <= w (rows without chunks share
// their successor's offset)
__device__ __forceinline__ int row_of(const RowPlan &R, int ns, int w) {
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (R.pre[mid] <= w)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ---- the staged replay: every pass serves all steps of one stage
__device__ __forceinline__ int flag_cand(int J) { return -(2 + 3 * J); }
__device__ __forceinline__ int flag_drop(int J) { return -(3 + 3 * J); }
__device__ __forceinline__ int flag_fresh(int J) { return -(4 + 3 * J); }

// pruneRedundancy of every step with a tree (the body of k_rep_redundancy,
// with the step's own block range standing for the grid)
__global__ void k_rep_red_multi(const RepOp *__restrict__ ops, const int *__restrict__ blk_off, int nops,
                                const Entry *__restrict__ E, float r, int *__restrict__ keepv) {
    __shared__ RowPlan R;
    __shared__ int wtmp[kBlock / 64];
    __shared__ int skip;
    __shared__ bool last;
    const int o = seg_of(blk_off, nops, blockIdx.x);
    const RepOp &Op = ops[o];
    const int nb = blk_off[o + 1] - blk_off[o], bi = blockIdx.x - blk_off[o];
    const int nlist = Op.nlist, nW = Op.nW, tree = Op.tree;
    const int *list = Op.list;
    int *rowflag = Op.rowflag;
    const double lhs = (double) ((float) nW * r);
    for (int seg = 0; seg < nlist; seg += kSegRows) {
        const int ns = min(kSegRows, nlist - seg);
        const int total = plan_rows(E, list, seg, ns, nW, R, wtmp);
        for (int w = bi; w < total; w += nb) {
            const int i = row_of(R, ns, w);
            const int h = seg + i;
            if (threadIdx.x == 0) skip = ld_relaxed(&rowflag[2 * h + 1]);
            __syncthreads();
            const bool sk = skip != 0;
            __syncthreads();
            if (sk) continue;
            const Entry e = E[list[h]];
            const int t0 = (w - R.pre[i]) * kChunk, t1 = min(R.len[i], t0 + kChunk);
            bool touch = false, small = false;
            if (R.by_mem[i]) {
                for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
                    const int q = e.mem[t];
                    if (Op.tstamp[q] == tree) {
                        touch = true;
                        small |= lhs > (double) e.csize[e.comp[q]];
                    }
                }
            } else {
                for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
                    const int K = e.comp[Op.W[t]];
                    if (K >= 0 && e.inres[K]) {
                        touch = true;
                        small |= lhs > (double) e.csize[K];
                    }
                }
            }
            touch = __syncthreads_or(touch);
            small = __syncthreads_or(small);
            if (threadIdx.x == 0) {
                if (touch && !ld_relaxed(&rowflag[2 * h])) atomicOr(&rowflag[2 * h], 1);
                if (small) atomicOr(&rowflag[2 * h + 1], 1);
            }
        }
        __syncthreads();
    }
    __threadfence();
    if (threadIdx.x == 0) last = atomicAdd(Op.done, 1u) == (unsigned) nb - 1;
    __syncthreads();
    if (last) {
        __threadfence();
        int drop = 0;
        for (int h = threadIdx.x; h < nlist; h += blockDim.x) {
            const int t = ld_relaxed(&rowflag[2 * h]), sm = ld_relaxed(&rowflag[2 * h + 1]);
            drop |= (t && !sm) ? 1 : 0;
        }
        drop = __syncthreads_or(drop);
        if (threadIdx.x == 0) keepv[tree] = drop ? 0 : 1;
    }
}

// candidates of every join step: the components (of at least min_size
// points) of the points of its kept trees; one list entry per component
__global__ void k_rep_cand_multi(const RepOp *__restrict__ ops, const int *__restrict__ blk_off, int nops,
                                 const int *__restrict__ keepv, int min_size, int n) {
    const int o = seg_of(blk_off, nops, blockIdx.x);
    const RepOp &Op = ops[o];
    const int i = (blockIdx.x - blk_off[o]) * blockDim.x + threadIdx.x;
    int K = -1;
    if (Op.kind == 0) {
        if (i < Op.nW && keepv[Op.tree]) K = Op.comp[Op.W[i]];
    } else if (i < n) {
        const int t = Op.tstamp[i];
        if (t >= Op.t_lo && t < Op.t_hi && keepv[t]) K = Op.comp[i];
    }
    const int cj = flag_cand(Op.J);
    bool want = K >= 0 && Op.csize[K] >= min_size && Op.flag[K] != cj;
    // one claim per distinct component in the wave
    unsigned long long todo = __ballot(want);
    while (todo) {
        const int leader = __ffsll((long long) todo) - 1;
        const int lk = __shfl(K, leader);
        todo &= ~__ballot(want && K == lk);
        if (__lane_id() == leader) {
            const int old = atomicExch(&Op.flag[lk], cj);
            if (old != cj) {
                const int at = atomicAdd(&Op.words[0], 1);
                if (at < Op.ccap) Op.clist[at] = lk;
            }
        }
    }
}

// a candidate K is dropped when one of its points lies in a result cluster of
// a superspace H larger than |K| * r
__global__ void k_rep_check_multi(const RepOp *__restrict__ ops, const int *__restrict__ blk_off, int nops,
                                  const Entry *__restrict__ E, float r) {
    __shared__ RowPlan R;
    __shared__ int wtmp[kBlock / 64];
    const int o = seg_of(blk_off, nops, blockIdx.x);
    const RepOp &Op = ops[o];
    if (ld_relaxed(&Op.words[0]) == 0) return;  // no candidate (uniform over the step's blocks)
    const int nb = blk_off[o + 1] - blk_off[o], bi = blockIdx.x - blk_off[o];
    const int cj = flag_cand(Op.J), dj = flag_drop(Op.J);
    for (int seg = 0; seg < Op.nlist; seg += kSegRows) {
        const int ns = min(kSegRows, Op.nlist - seg);
        const int total = plan_rows(E, Op.list, seg, ns, -1, R, wtmp);
        for (int w = bi; w < total; w += nb) {
            const int i = row_of(R, ns, w);
            const Entry e = E[Op.list[seg + i]];
            const int t0 = (w - R.pre[i]) * kChunk, t1 = min(R.len[i], t0 + kChunk);
            for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
                const int q = e.mem[t];
                const int K = Op.comp[q];
                if (K < 0 || Op.flag[K] != cj) continue;
                if ((float) Op.csize[K] * r < (float) e.csize[e.comp[q]]) Op.flag[K] = dj;
            }
        }
        __syncthreads();
    }
}

// the surviving candidates of every join step join the result of its
// subspace (one block per step)
__global__ void k_rep_apply_multi(const RepOp *__restrict__ ops, Entry *__restrict__ E) {
    __shared__ int wtmp[kBlock / 64];
    const RepOp &Op = ops[blockIdx.x];
    if (Op.kind == 1) return;
    const int nc = min(Op.words[0], Op.ccap);
    const int cj = flag_cand(Op.J), fj = flag_fresh(Op.J);
    int added = 0;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) {
        const int K = Op.clist[k];
        if (Op.flag[K] == cj && !Op.inres[K]) {
            Op.inres[K] = 1;
            Op.flag[K] = fj;
            ++added;
        }
    }
    for (int o = 32; o > 0; o >>= 1) added += __shfl_down(added, o);
    if (__lane_id() == 0) wtmp[threadIdx.x / 64] = added;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < kBlock / 64; ++w) tot += wtmp[w];
        if (tot) E[Op.entry].nkept += tot;
        Op.words[1] = tot;
    }
}

constexpr int kMemPerBlock = 1024;

// members of the newly joined clusters of every join step
__global__ void k_rep_members_multi(const RepOp *__restrict__ ops, const int *__restrict__ blk_off, int nops,
                                    Entry *__restrict__ E, int n) {
    __shared__ int wcount[kBlock / 64];
    __shared__ int base;
    const int o = seg_of(blk_off, nops, blockIdx.x);
    const RepOp &Op = ops[o];
    if (Op.words[1] == 0) return;  // nothing joined (uniform over the step's blocks)
    const int fj = flag_fresh(Op.J);
    constexpr int PER = kMemPerBlock / kBlock;
    const int i0 = (blockIdx.x - blk_off[o]) * kMemPerBlock;
    bool in[PER];
    int mine = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int i = i0 + u * kBlock + threadIdx.x;
        bool f = false;
        if (i < n) {
            const int K = Op.comp[i];
            f = K >= 0 && Op.flag[K] == fj;
        }
        in[u] = f;
        mine += f ? 1 : 0;
    }
    int incl = mine;
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d);
        if (__lane_id() >= d) incl += v;
    }
    if (__lane_id() == 63) wcount[threadIdx.x / 64] = incl;
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < (int) threadIdx.x / 64) before += wcount[w];
        total += wcount[w];
    }
    if (threadIdx.x == 0) base = total > 0 ? atomicAdd(&E[Op.entry].nmem, total) : 0;
    __syncthreads();
    int slot = base + before + incl - mine;
    int *mem = E[Op.entry].mem;
#pragma unroll
    for (int u = 0; u < PER; ++u)
        if (in[u]) mem[slot++] = i0 + u * kBlock + threadIdx.x;
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
    fill_span(p, v, count);
}

}  // namespace

#define LAUNCH(kernel, grid, block, lds, st, ...) hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__)

static_assert(kBlockThreads == kBlock && kQueriesPerBlock == QPB, "launch geometry constants disagree");

static unsigned query_blocks_n(int n) { return (unsigned) ((n + QPB - 1) / QPB); }
static size_t count_lds(int kmax) {
    return (sizeof(float) + sizeof(int)) * QPB * (size_t) (kmax > 0 ? kmax : 1) + 8 * QPB +
           sizeof(int) * QPB * kRangeWords;
}

hipError_t launch_b_flags(const FlagTree *T, const int *blk_off, int nt, int nblk, const uint8_t *cellC, int n,
                          int cells, int *flags, hipStream_t st) {
    if (nblk <= 0) return hipSuccess;
    LAUNCH(k_b_flags, dim3(nblk), dim3(kBlock), 0, st, T, blk_off, nt, cellC, n, cells, flags);
    return hipGetLastError();
}

hipError_t launch_b_index(const DevData &D, const BatchSub *subs, int nsub, const BatchIndex &X, bool need_es,
                          const int *miss_reps, int nmiss, int hmiss, const SortCopy *stores, int nstore,
                          const SortCopy *hits, int nhit, int max_cell, const ColumnKeep &cols, hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    // the scan over the sorted indexes' cells ends with their total, the
    // first position of the kept ones
    const int nscan = hmiss + 1;
    LAUNCH(k_b_fill<int>, dim3(fill_blocks((size_t) nscan, sizeof(int))), dim3(kBlock), 0, st, X.hist, 0, (size_t) nscan);
    if (nmiss > 0)
        LAUNCH(k_b_keys, dim3(blocks_for(D.n), nmiss), dim3(kBlock), 0, st, D, subs, miss_reps, X.keys, X.rank,
               X.hist);
    hipError_t e = launch_exclusive_sum(X.scan_temp, X.scan_bytes, X.hist, X.off, nscan, st);
    if (e != hipSuccess) return e;
    if (nmiss > 0)
        LAUNCH(k_b_scatter, dim3(blocks_for(D.n), nmiss), dim3(kBlock), 0, st, D, subs, miss_reps, X.keys, X.rank,
               X.off, X.order);
    const unsigned span = blocks_for((size_t) std::max(D.n, max_cell + 1));
    if (nstore > 0)
        LAUNCH(k_b_sort_out, dim3(span, nstore), dim3(kBlock), 0, st, stores, D.n, X.order, X.keys, X.off);
    if (nhit > 0) LAUNCH(k_b_sort_in, dim3(span, nhit), dim3(kBlock), 0, st, hits, D.n, X.order, X.keys, X.off);
    LAUNCH(k_b_gather, dim3(blocks_for(D.n), X.nrep), dim3(kBlock), 0, st, D, subs, X.reps, X.order, X.Ys, X.Es,
           X.sigS, (int) need_es, cols.src, cols.store, cols.kmax);
    LAUNCH(k_b_gather_own, dim3(blocks_for(D.n), nsub), dim3(kBlock), 0, st, D, subs, X.order, X.Ys, X.Es, X.sigS,
           X.sig, (int) need_es, cols.own_src, cols.own_store);
    return hipGetLastError();
}

// Ordered compaction of the query flags (one byte per sorted position, 0 or
// 1) into the positions that are set: a count per tile of 4,096 flags, an
// exclusive scan of the tile counts, and a write pass that ranks each lane's
// 16 flags inside its tile.  Every lane reads its flags as one 16-byte load;
// the flag array is padded to a multiple of 16 bytes and bytes at or past nb
// are ignored.
constexpr int kFlagsPerLane = 16;
constexpr int kFlagTile = kBlock * kFlagsPerLane;

__device__ __forceinline__ int set_bytes(uint32_t w) {
    w |= w >> 4;  // bit 0 of each byte: the OR of the byte's bits
    w |= w >> 2;
    w |= w >> 1;
    return __builtin_popcount(w & 0x01010101u);
}

__device__ __forceinline__ uint4 lane_flags(const uint8_t *__restrict__ qflag, size_t at, size_t nb) {
    if (at >= nb) return make_uint4(0, 0, 0, 0);
    uint4 v = *reinterpret_cast<const uint4 *>(qflag + at);
    if (at + kFlagsPerLane > nb) {
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
        for (int q = 0; q < 4; ++q)
            for (int c = 0; c < 4; ++c)
                if (at + 4 * q + c >= nb) w[q] &= ~(0xffu << (8 * c));
        v = make_uint4(w[0], w[1], w[2], w[3]);
    }
    return v;
}

__global__ void __launch_bounds__(kBlock) k_b_flag_count(const uint8_t *__restrict__ qflag, size_t nb, int ntile,
                                                          int *__restrict__ tcnt) {
    __shared__ int wsum[kBlock / 64];
    const size_t at = ((size_t) blockIdx.x * kBlock + threadIdx.x) * kFlagsPerLane;
    const uint4 v = lane_flags(qflag, at, nb);
    int c = set_bytes(v.x) + set_bytes(v.y) + set_bytes(v.z) + set_bytes(v.w);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (__lane_id() == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int q = 0; q < kBlock / 64; ++q) t += wsum[q];
        tcnt[blockIdx.x] = t;
        if (blockIdx.x == 0) tcnt[ntile] = 0;  // the scan's last entry is the total
    }
}

__global__ void __launch_bounds__(kBlock) k_b_flag_write(const uint8_t *__restrict__ qflag, size_t nb,
                                                          const int *__restrict__ toff, int ntile,
                                                          int *__restrict__ qlist, int *__restrict__ nsel) {
    __shared__ int wsum[kBlock / 64];
    const size_t at = ((size_t) blockIdx.x * kBlock + threadIdx.x) * kFlagsPerLane;
    const uint4 v = lane_flags(qflag, at, nb);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    int m[4], c = 0;
    for (int q = 0; q < 4; ++q) {
        uint32_t x = w[q];
        x |= x >> 4;
        x |= x >> 2;
        x |= x >> 1;
        m[q] = (int) (x & 0x01010101u);
        c += __builtin_popcount((uint32_t) m[q]);
    }
    const int lane = __lane_id();
    int x = c;  // inclusive scan over the wave
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    int o = toff[blockIdx.x] + x - c;
    for (int q = 0; q < (int) (threadIdx.x >> 6); ++q) o += wsum[q];
    for (int q = 0; q < 4; ++q) {
        uint32_t mm = (uint32_t) m[q];
        while (mm) {
            const int bit = __builtin_ctz(mm);
            qlist[o++] = (int) (at + 4 * q + (bit >> 3));
            mm &= mm - 1;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *nsel = toff[ntile];
}

static int flag_tiles(size_t items) { return (int) ((items + kFlagTile - 1) / kFlagTile); }

static size_t align256(size_t b) { return (b + 255) & ~(size_t) 255; }

size_t select_temp_bytes(int items) {
    const int nt = flag_tiles((size_t) std::max(items, 1));
    return 2 * align256((size_t) (nt + 1) * sizeof(int)) + align256(scan_temp_bytes(nt + 1));
}

hipError_t launch_flag_compact(const uint8_t *flags, size_t nb, void *temp, size_t temp_bytes, int *out, int *nsel,
                               hipStream_t st) {
    const int nt = flag_tiles(nb);
    const size_t cnt_bytes = align256((size_t) (nt + 1) * sizeof(int));
    char *base = static_cast<char *>(temp);
    int *tcnt = reinterpret_cast<int *>(base);
    int *toff = reinterpret_cast<int *>(base + cnt_bytes);
    LAUNCH(k_b_flag_count, dim3(nt), dim3(kBlock), 0, st, flags, nb, nt, tcnt);
    hipError_t e = launch_exclusive_sum(base + 2 * cnt_bytes, temp_bytes - 2 * cnt_bytes, tcnt, toff, nt + 1, st);
    if (e != hipSuccess) return e;
    LAUNCH(k_b_flag_write, dim3(nt), dim3(kBlock), 0, st, flags, nb, toff, nt, out, nsel);
    return hipGetLastError();
}

hipError_t launch_b_queries(const BatchSub *subs, int nsub, int n, const QuerySeg *segs, const int *blk_off,
                            int nseg, int nblk, const BatchIndex &X, hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    const size_t nb = (size_t) nsub * n;
    LAUNCH(k_b_fill<uint8_t>, dim3(fill_blocks(nb, 1)), dim3(kBlock), 0, st, X.qflag, (uint8_t) 0, nb);
    if (nblk > 0)
        LAUNCH(k_b_qmark, dim3(nblk), dim3(kBlock), 0, st, subs, segs, blk_off, nseg, X.keys, X.qflag);
    hipError_t e = launch_flag_compact(X.qflag, nb, X.select_temp, X.select_bytes, X.qlist, X.nsel, st);
    if (e != hipSuccess) return e;
    LAUNCH(k_b_qstart, dim3(blocks_for(nsub + 1)), dim3(kBlock), 0, st, X.qlist, X.nsel, nsub, n, X.qstart);
    return hipGetLastError();
}

template <int K>
static void launch_count_k(const BatchSub *subs, int nsub, const int *blk_off, int nblk, int n, size_t lds,
                           const BatchIndex &X, float s_star, float eps, double phirel, bool phi,
                           unsigned long long *stat, hipStream_t st) {
    if (blk_off) {
        if (phi)
            LAUNCH((k_b_count<true, true, K>), dim3(nblk), dim3(kBlock), lds, st, subs, blk_off, nsub, n, X.off,
                   X.order, X.Ys, X.Es, X.sig, X.qlist, X.qstart, s_star, eps, phirel, X.nb, X.fb, X.nfb, X.fbcap, stat);
        else
            LAUNCH((k_b_count<true, false, K>), dim3(nblk), dim3(kBlock), lds, st, subs, blk_off, nsub, n, X.off,
                   X.order, X.Ys, X.Es, X.sig, X.qlist, X.qstart, s_star, eps, phirel, X.nb, X.fb, X.nfb, X.fbcap, stat);
    } else {
        const dim3 grid(query_blocks_n(n), nsub);
        if (phi)
            LAUNCH((k_b_count<false, true, K>), grid, dim3(kBlock), lds, st, subs, (const int *) nullptr, nsub, n,
                   X.off, X.order, X.Ys, X.Es, X.sig, X.qlist, X.qstart, s_star, eps, phirel, X.nb, X.fb, X.nfb,
                   X.fbcap, stat);
        else
            LAUNCH((k_b_count<false, false, K>), grid, dim3(kBlock), lds, st, subs, (const int *) nullptr, nsub, n,
                   X.off, X.order, X.Ys, X.Es, X.sig, X.qlist, X.qstart, s_star, eps, phirel, X.nb, X.fb, X.nfb,
                   X.fbcap, stat);
    }
}

hipError_t launch_phi_exact(const DevData &D, const BatchSub *sub, const int2 *fb, const int *nfb, int fbcap,
                            const int *rank_order, float eps, float s_star, hipStream_t st) {
    LAUNCH(k_b_phi_exact, dim3(1024), dim3(kBlock), 0, st, D, sub, fb, nfb, fbcap, rank_order, eps, s_star);
    return hipGetLastError();
}

hipError_t launch_b_count(const BatchSub *subs, int nsub, const int *blk_off, int nblk, int n, int kmax, int kfast,
                          const BatchIndex &X, float s_star, float eps, bool phi, bool phi_all_exact,
                          const DevData &D, const int *rank_order, unsigned long long *stat, const KernelTimer &tm,
                          hipStream_t st) {
    // phi_all_exact: every point goes through the leaf-order recomputation (a test hook)
    const double phirel = phi_all_exact ? 1e30 : kPhiRelDefault;
    if (nsub <= 0) return hipSuccess;
    if (blk_off && nblk <= 0) return hipSuccess;
    if (phi) LAUNCH(k_b_fill<int>, dim3(1), dim3(kBlock), 0, st, X.nfb, 0, (size_t) 1);
    if (tm.start) (void) hipEventRecord(tm.start, st);
    const size_t lds = count_lds(kmax);
    switch (kfast) {
    case 1: launch_count_k<1>(subs, nsub, blk_off, nblk, n, lds, X, s_star, eps, phirel, phi, stat, st); break;
    case 2: launch_count_k<2>(subs, nsub, blk_off, nblk, n, lds, X, s_star, eps, phirel, phi, stat, st); break;
    case 3: launch_count_k<3>(subs, nsub, blk_off, nblk, n, lds, X, s_star, eps, phirel, phi, stat, st); break;
    case 4: launch_count_k<4>(subs, nsub, blk_off, nblk, n, lds, X, s_star, eps, phirel, phi, stat, st); break;
    case 5: launch_count_k<5>(subs, nsub, blk_off, nblk, n, lds, X, s_star, eps, phirel, phi, stat, st); break;
    case 6: launch_count_k<6>(subs, nsub, blk_off, nblk, n, lds, X, s_star, eps, phirel, phi, stat, st); break;
    case 7: launch_count_k<7>(subs, nsub, blk_off, nblk, n, lds, X, s_star, eps, phirel, phi, stat, st); break;
    default: launch_count_k<0>(subs, nsub, blk_off, nblk, n, lds, X, s_star, eps, phirel, phi, stat, st); break;
    }
    if (tm.stop) (void) hipEventRecord(tm.stop, st);
    if (phi)
        LAUNCH(k_b_phi_exact, dim3(1024), dim3(kBlock), 0, st, D, subs, X.fb, X.nfb, X.fbcap, rank_order, eps, s_star);
    return hipGetLastError();
}

hipError_t launch_b_components(const BatchSub *subs, int nsub, int n, int kmax, int kfast, const BatchIndex &X,
                               float s_star, hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    LAUNCH(k_b_dense_init, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.order, X.dsorted, X.par);
    for (int round = 0; round < kSampleEdges; ++round) {
        LAUNCH(k_b_link_sample, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.nb, round, X.par);
        LAUNCH(k_b_uf_compress, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.par);
    }
    LAUNCH(k_b_giant, dim3(nsub), dim3(256), 0, st, subs, n, X.par, X.giant);
    const dim3 grid(query_blocks_n(n), nsub);
    const size_t lds = count_lds(kmax);
#define LINK_FULL(K)                                                                                                   \
    LAUNCH(k_b_link_full<K>, grid, dim3(kBlock), lds, st, subs, n, X.off, X.order, X.Ys, X.Es, X.sig, X.dsorted,      \
           s_star, X.giant, X.par)
    switch (kfast) {
    case 1: LINK_FULL(1); break;
    case 2: LINK_FULL(2); break;
    case 3: LINK_FULL(3); break;
    case 4: LINK_FULL(4); break;
    case 5: LINK_FULL(5); break;
    case 6: LINK_FULL(6); break;
    case 7: LINK_FULL(7); break;
    default: LINK_FULL(0); break;
    }
#undef LINK_FULL
    LAUNCH(k_b_uf_flatten, dim3(blocks_for(n), nsub), dim3(kBlock), 0, st, subs, n, X.par);
    return hipGetLastError();
}

hipError_t launch_b_partition(const PartPair *pairs, const int *blk_off, int npairs, int nblk, const uint8_t *cellC,
                              int n, int thr, int wbit, int min_size, hipStream_t st) {
    if (nblk <= 0) return hipSuccess;
    LAUNCH(k_b_part_count, dim3(nblk), dim3(kBlock), 0, st, pairs, blk_off, npairs, cellC, n, thr, wbit);
    LAUNCH(k_b_part_scatter, dim3(nblk), dim3(kBlock), 0, st, pairs, blk_off, npairs, cellC, n, thr, wbit, min_size);
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

int rep_red_blocks(int nlist, int nW) { return std::max(1, std::min(nlist + nW / kChunk, 256)); }
int rep_check_blocks(int nlist, int n) { return nlist > 0 ? std::min(nlist + n / kChunk, 256) : 0; }
int rep_mem_blocks(int n) { return (n + kMemPerBlock - 1) / kMemPerBlock; }

hipError_t launch_rep_stage(const RepOp *ops, int nops, const int *blk_red, int nblk_red, const int *blk_cand,
                            int nblk_cand, const int *blk_chk, int nblk_chk, const int *blk_mem, int nblk_mem,
                            Entry *E, float r, int min_size, int n, int *keepv, hipStream_t st) {
    if (nops <= 0) return hipSuccess;
    if (nblk_red > 0) LAUNCH(k_rep_red_multi, dim3(nblk_red), dim3(kBlock), 0, st, ops, blk_red, nops, E, r, keepv);
    if (nblk_cand > 0)
        LAUNCH(k_rep_cand_multi, dim3(nblk_cand), dim3(kBlock), 0, st, ops, blk_cand, nops, keepv, min_size, n);
    if (nblk_chk > 0) LAUNCH(k_rep_check_multi, dim3(nblk_chk), dim3(kBlock), 0, st, ops, blk_chk, nops, E, r);
    LAUNCH(k_rep_apply_multi, dim3(nops), dim3(kBlock), 0, st, ops, E);
    if (nblk_mem > 0) LAUNCH(k_rep_members_multi, dim3(nblk_mem), dim3(kBlock), 0, st, ops, blk_mem, nops, E, n);
    return hipGetLastError();
}

hipError_t launch_b_entries(Entry *E, int at, const Entry *src, int count, int *nE, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    LAUNCH(k_b_entries, dim3(1), dim3(kBlock), 0, st, E, at, src, count, nE);
    return hipGetLastError();
}

}  // namespace inscy
