// Device helpers shared by the kernel files of the MI355X INSCY engine:
// wave-level compaction, the grid-index candidate walk and distance test,
// union-find, and the hash of the path prune.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "inscy_internal.h"

namespace inscy {


constexpr int kBlock = 256;
constexpr int kLanesPerQuery = 16;  // lanes cooperating on one query's candidate ranges

__device__ __forceinline__ int ld_relaxed(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_relaxed(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// Wave-aggregated slot allocation for order-free stream compaction: one
// atomic per wave.  Must be reached by every active lane of the wave.
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

// counter[key] += (number of lanes with that key and want), one atomic per
// distinct key in the wave; returns each wanting lane's slot (old value +
// rank among the lanes sharing its key).  Must be reached by every active lane.
__device__ __forceinline__ int wave_add_by_key(int *counter, int key, bool want) {
    unsigned long long todo = __ballot(want);
    const int lane = __lane_id();
    int result = 0;
    while (todo) {
        int leader = __ffsll((long long) todo) - 1;
        int lk = __shfl(key, leader);
        unsigned long long same = __ballot(want && key == lk) & todo;
        int b = 0;
        if (lane == leader) b = atomicAdd(&counter[lk], __popcll(same));
        b = __shfl(b, leader);
        if ((same >> lane) & 1ull) result = b + __popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return result;
}

inline unsigned blocks_for(size_t count) { return (unsigned) ((count + kBlock - 1) / kBlock); }

template <class Range>
__device__ __forceinline__ void for_ranges(const GridArg &G, const int *__restrict__ off, const int *qe, int stride,
                                           Range f) {
    const int C = G.C;
    if (G.j == 1) {
        int c0 = qe[G.tI[0] * stride];
        f(off[max(c0 - 1, 0)], off[min(c0 + 1, C - 1) + 1]);
    } else if (G.j == 2) {
        int c0 = qe[G.tI[0] * stride], c1 = qe[G.tI[1] * stride];
        int lo = max(c1 - 1, 0), hi = min(c1 + 1, C - 1);
        for (int x = max(c0 - 1, 0); x <= min(c0 + 1, C - 1); ++x) f(off[x * C + lo], off[x * C + hi + 1]);
    } else {
        int c0 = qe[G.tI[0] * stride], c1 = qe[G.tI[1] * stride], c2 = qe[G.tI[2] * stride];
        int lo = max(c2 - 1, 0), hi = min(c2 + 1, C - 1);
        for (int x = max(c0 - 1, 0); x <= min(c0 + 1, C - 1); ++x)
            for (int y = max(c1 - 1, 0); y <= min(c1 + 1, C - 1); ++y) {
                int base = (x * C + y) * C;
                f(off[base + lo], off[base + hi + 1]);
            }
    }
}

// candidate at sorted position jj against the staged query; S order, no FMA
__device__ __forceinline__ bool is_neighbour(const SubspaceArg &S, const GridArg &G, int n,
                                             const float *__restrict__ Ys, const uint16_t *__restrict__ Es,
                                             const float *qy, const int *qe, int stride, int jj, float s_star) {
    float acc = 0.0f;
    for (int t = 0; t < S.k; ++t) {
        if ((G.checkmask >> t) & 1ull) {
            int de = (int) Es[(size_t) t * n + jj] - qe[t * stride];
            if (de < -1 || de > 1) return false;
        }
        float diff = qy[t * stride] - Ys[(size_t) t * n + jj];
        float sq = diff * diff;
        acc = acc + sq;
        if (acc > s_star) return false;
    }
    return true;
}

__device__ __forceinline__ void stage_query(const SubspaceArg &S, int n, const float *__restrict__ Ys,
                                            const uint16_t *__restrict__ Es, int i, float *sy, int *se, int stride) {
    for (int t = 0; t < S.k; ++t) {
        sy[t * stride] = Ys[(size_t) t * n + i];
        se[t * stride] = Es[(size_t) t * n + i];
    }
}

// the ceil(1/eps)-grid cell of a coordinate, exactly as k_discretise computes
// it (correctly rounded 1/ecells and quotient): cells recomputed from a
// gathered coordinate equal the stored ones
__device__ __forceinline__ float ecell_width(int ecells) { return __fdiv_rn(1.0f, (float) ecells); }
__device__ __forceinline__ int ecell_of(float x, float ecs, int ecells) { return min((int) __fdiv_rn(x, ecs), ecells - 1); }

// p[0, count) = v with 16-byte stores over the aligned middle (grid-stride;
// the ends one element at a time)
template <class T>
__device__ __forceinline__ void fill_span(T *p, T v, size_t count) {
    static_assert(16 % sizeof(T) == 0, "element size divides 16");
    constexpr size_t per = 16 / sizeof(T);
    size_t head = (per - ((size_t) reinterpret_cast<uintptr_t>(p) / sizeof(T)) % per) % per;
    if (head > count) head = count;
    const size_t nvec = (count - head) / per, tail = head + nvec * per;
    union {
        T e[per];
        uint4 u;
    } w;
#pragma unroll
    for (size_t k = 0; k < per; ++k) w.e[k] = v;
    const size_t t0 = (size_t) blockIdx.x * blockDim.x + threadIdx.x, step = (size_t) gridDim.x * blockDim.x;
    uint4 *q = reinterpret_cast<uint4 *>(p + head);
    for (size_t j = t0; j < nvec; j += step) q[j] = w.u;
    if (t0 < head) p[t0] = v;
    if (tail + t0 < count) p[tail + t0] = v;
}
inline unsigned fill_blocks(size_t count, size_t elem) {
    const size_t vec = (count * elem + 15) / 16 + 2;
    return (unsigned) std::max<size_t>(1, std::min<size_t>((vec + 255) / 256, 4096));
}

// Union-find over parents that only move towards smaller ids (par[x] <= x,
// x a root iff par[x] == x; unions link the larger root under the smaller
// with a device-scope compare-and-swap).  Every value ever held by par[x] is
// an ancestor of x, so the finds use plain loads and stores: a copy that an
// XCD's L2 still holds after another XCD changed the word yields an older
// ancestor, never a wrong set.  uf_union retries with the value its failed
// swap returns (always current, and smaller), so it ends; a find with no
// union running beside it (compression, flattening) returns the current root.
__device__ __forceinline__ int uf_find(int *par, int x) {
    int p = par[x];
    while (p != x) {
        int g = par[p];
        if (g != p) par[x] = g;  // path halving: g is an ancestor of x
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
        int old = atomicCAS(&par[b], b, a);
        if (old == b) return;
        b = old;
    }
}

// OR of every lane's m into *dst (all lanes of the wave take part)
__device__ __forceinline__ void wave_or_mask(dmask_t m, dmask_t *dst) {
    uint64_t lo = dlo(m), hi = dhi(m);
    for (int off = 32; off > 0; off >>= 1) {
        lo |= __shfl_xor(lo, off);
        hi |= __shfl_xor(hi, off);
    }
    if (__lane_id() == 0) {
        unsigned long long *w = reinterpret_cast<unsigned long long *>(dst);
        if (lo) atomicOr(&w[0], (unsigned long long) lo);
        if (hi) atomicOr(&w[1], (unsigned long long) hi);
    }
}

// bucket of coordinate x in an index's last dimension with F buckets per eps
// cell (C cells): floor(x C F) within x's own cell's F buckets, so that the
// order of buckets follows the order of cells and of coordinates (monotone in x)
__device__ __forceinline__ int fine_key(float x, int C, int F) {
    x = fminf(fmaxf(x, 0.0f), 1.0f);
    const int c = ecell_of(x, ecell_width(C), C);
    const int f = (int) floor((double) x * (double) C * (double) F);
    return min(max(f, c * F), c * F + F - 1);
}

// density tests on a stored value: counts (bit == 0) or phi decision bits
// (bit 1 weak-dense, bit 2 dense)
__device__ __forceinline__ bool meets(int v, int thr, int bit) { return bit ? (v >= 0 && (v & bit) != 0) : v >= thr; }

// term of phi (ClusteringCpu.cu:116-124) for a neighbour at squared distance
// acc: d = sqrtf(acc) / eps in float, then 1 - d*d in double
__device__ __forceinline__ double phi_term(float acc, float eps) {
    const float dist = __fsqrt_rn(acc);
    const float r = __fdiv_rn(dist, eps);
    const float sq = __fmul_rn(r, r);
    return 1.0 - (double) sq;
}

// The reference sums phi in float in the order of the neighbourhood tree's
// leaves.  The passes here sum the same terms in double in scan order: with N
// terms the float sum in leaf order lies within a relative N * 2^-22 of it
// (each float rounding moves the running sum by at most 2^-24 of the final
// value), so a threshold outside that band is decided exactly; points inside
// it are recomputed in leaf order (k_b_phi_exact).
constexpr double kPhiRelDefault = 1.0 / 4194304.0;

// Stored value of phi mode: (q << 2) | bits, bits the decisions (1
// weak-dense, 2 dense) and q = ceil(4 U) for an upper bound U on the
// reference's phi of the point in S and in every superset of S (q =
// kPhiNoBound: none known).  phi is monotone like the counts: a superset of S
// has fewer neighbours, each term no larger (the float distance sums over
// more dimensions, every rounding monotone), so T, the exact sum of the
// float terms, only falls; the float leaf-order sum of N terms lies within a
// relative N 2^-22 of T (the band of the passes), hence U = T_upper (1 + N
// 2^-22) with N the neighbours in S bounds phi in every superset.
constexpr int kPhiNoBound = (1 << 29) - 1;
__device__ __forceinline__ int phi_code(int bits, double ub) {
    const int q = ub >= 0.0 && ub < (double) (kPhiNoBound - 1) * 0.25 ? (int) ceil(ub * 4.0) : kPhiNoBound;
    return (q << 2) | bits;
}

constexpr uint64_t kEmpty = ~0ull;

__device__ __forceinline__ uint64_t mix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

// a and b share their c-grid cells on every unrestricted dimension at SCY-tree
// levels 0..upto (the regular-node prefix of the tree)
__device__ __forceinline__ bool same_prefix(const DevData &D, dmask_t smask, int a, int b, int upto) {
    if (b < 0 || b >= D.n) return false;
    for (int t = 0; t <= upto; ++t) {
        const int k = D.level[t];
        if (dhas(smask, k)) continue;
        if (D.cellC[(size_t) k * D.n + a] != D.cellC[(size_t) k * D.n + b]) return false;
    }
    return true;
}

}  // namespace inscy
