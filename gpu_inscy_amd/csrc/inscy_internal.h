// Internal declarations shared by the HIP kernels and the host driver of the
// MI355X INSCY engine.  Not part of the public ABI (include/inscy.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

namespace inscy {

// Sets of dimensions (subspaces, s-connection paths, the near-edge marks of a
// point) are 128-bit masks, bit j for dimension j; their numeric order is the
// order of the reference's vec_cmp (util.cu:633-653: the larger highest
// differing dimension is the larger set).
typedef unsigned __int128 dmask_t;
constexpr int kMaxDims = 128;     // d
constexpr int kMaxSubDims = 64;   // dimensions of one subspace (positions in S are uint64 masks)

__host__ __device__ __forceinline__ dmask_t dbit(int j) { return (dmask_t) 1 << j; }
__host__ __device__ __forceinline__ bool dhas(dmask_t m, int j) { return (bool) ((m >> j) & 1u); }
// dimensions above j (j in [-1, kMaxDims))
__host__ __device__ __forceinline__ dmask_t dabove(int j) {
    return j + 1 >= kMaxDims ? (dmask_t) 0 : ~(dmask_t) 0 << (j + 1);
}
__host__ __device__ __forceinline__ uint64_t dlo(dmask_t m) { return (uint64_t) m; }
__host__ __device__ __forceinline__ uint64_t dhi(dmask_t m) { return (uint64_t) (m >> 64); }
__host__ __device__ __forceinline__ dmask_t dmake(uint64_t lo, uint64_t hi) { return (dmask_t) hi << 64 | lo; }
__host__ __device__ __forceinline__ int dcount(dmask_t m) {
    return __builtin_popcountll(dlo(m)) + __builtin_popcountll(dhi(m));
}
// lowest dimension of a non-empty mask
__host__ __device__ __forceinline__ int dfirst(dmask_t m) {
    return dlo(m) ? __builtin_ctzll(dlo(m)) : 64 + __builtin_ctzll(dhi(m));
}
static_assert(sizeof(dmask_t) == 16 && alignof(dmask_t) == 16, "one layout on host and device");
// Arrays of s-connection path masks hold mw 64-bit words per entry: 1 when
// d <= 64 (half the bytes of the partitions and path lists), 2 above.
__host__ __device__ __forceinline__ dmask_t mload(const uint64_t *a, long long i, int mw) {
    return mw == 1 ? (dmask_t) a[i] : reinterpret_cast<const dmask_t *>(a)[i];
}
__host__ __device__ __forceinline__ void mstore(uint64_t *a, long long i, dmask_t v, int mw) {
    if (mw == 1)
        a[i] = dlo(v);
    else
        reinterpret_cast<dmask_t *>(a)[i] = v;
}

struct DMaskHash {
    size_t operator()(dmask_t m) const {
        uint64_t h = dlo(m) * 0x9E3779B97F4A7C15ull ^ (dhi(m) + 0x632BE59BD9B4E019ull);
        return (size_t) (h ^ (h >> 29));
    }
};

// Subspace argument passed by value to kernels: k ascending dimension ids.
struct SubspaceArg {
    int k;
    int dims[kMaxSubDims];
};

// Device-resident discretised data set.  All per-dimension arrays are
// dimension-major ([d][n]) so that a wave reading one dimension of 64
// consecutive points issues one coalesced load.
struct DevData {
    const float *Xt;        // [d][n] coordinates
    const uint8_t *cellC;   // [d][n] cell in the c-grid (ScyTreeNode.cu:100-102)
    const uint16_t *cellE;  // [d][n] cell in the ceil(1/eps)-grid (inscy_map.cpp:243-244)
    const int *level;       // [d] dimension at each SCY-tree level (get_subspace_order, inscy_map.cpp:61-121)
    const int *lvl_of;      // [d] level of each dimension
    int n, d, cells, ecells;
};

// Grid index of one subspace: the points are sorted by their eps-grid cells in
// j (1..3) of the subspace's dimensions.
struct GridArg {
    int j;               // number of index dimensions
    int C;               // cells per index dimension (ceil(1/eps))
    int tI[3];           // positions within the subspace of the index dimensions
    int F;               // buckets per cell of the last index dimension (0 or 1: the cells themselves)
    uint64_t checkmask;  // bit t: dimension S.dims[t] needs the per-candidate cell test
};

// grid cells of an index: C^j, the last index dimension in F buckets per cell
inline long long grid_cells(const GridArg &G) {
    long long c = G.F > 1 ? G.F : 1;
    for (int a = 0; a < G.j; ++a) c *= G.C;
    return c;
}
constexpr int kFineBuckets = 4;  // buckets per cell of a batched index's last dimension (indexes of j <= 2)

// Scratch of one subspace grid pass (device pointers, sized for n points,
// k dims and C^j + 1 grid cells).
struct GridBuffers {
    int *keys;        // [n]
    int *rank;        // [n] position of a point within its grid cell
    int *hist;        // [ncell+1]
    int *off;         // [ncell+1] cell starts
    int *order;       // [n] point id at each sorted position
    float *Ys;        // [k][n] coordinates in sorted order
    uint16_t *Es;     // [k][n] eps-grid cells in sorted order
    uint8_t *dsorted; // [n] dense flag per sorted position
    int *par;         // [n] union-find parents
    int *giant;       // [1] root of the largest sampled component
    void *scan_temp;
    size_t scan_bytes;
};

// One subspace of the device-resident result table: its dense components and
// which of them the result holds (nkept > 0 <=> the subspace is in the result).
struct Entry {
    dmask_t mask;
    const int *comp;   // [n] component root of each point, -1 if not dense
    const int *csize;  // [n] component sizes indexed by root
    uint8_t *inres;    // [n] component kept in the result
    int *mem;          // [n] points of the kept components, nmem of them
    int nkept;
    int nmem;
};

// Optional event pair bracketing the neighbour-count kernel of a launch.
struct KernelTimer {
    hipEvent_t start = nullptr, stop = nullptr;
};

// ------------------------------------------------------------ level-synchronous passes
// Every pass of the level-synchronous driver runs over many segments (trees,
// subspaces, (tree, dimension) pairs) in one launch.  Segment s owns blocks
// [blk_off[s], blk_off[s+1]) of the grid; the descriptors live in device memory.
//
// Path origins carry one extra bit: x >= 0 when the origin is a point of the
// tree that holds the path, ~x when it is not (the s-connection of a point
// pruned by the weak-density test).

// One subspace of a batch.  Its per-subspace state (query flags, dense flags,
// packed cells, sample edges, union-find parents) lives at [base, base + n).
// Its grid index (keys, ranks, sorted order) lives at [obase, obase + n) with
// grid cells at [hoff, hoff + ncell); the children S + {m} of one frontier
// mask S (m is always the last dimension) share one index, built by the first
// of them (rep = 1), and the coordinates of S's dimensions in sorted order.
// Column t of the sorted coordinates and cells starts at mcoord for t == pos_m
// (the child's own dimension), else at coord + t*n; pos_m = k when nothing is
// shared.
constexpr int kMaxBounds = 8;
struct BatchSub {
    SubspaceArg S;
    GridArg G;
    long long base, coord, hoff, qoff;
    long long obase, mcoord;
    int pos_m, rep;
    int *cnt;           // [n] neighbour counts (capped), -1 = not counted
    const int *bound[kMaxBounds];  // [n] complete counts of subspaces S minus one dimension
    int *comp;          // [n] component roots (components pass)
    int *csize;         // [n] component sizes (components pass)
    int nbound, ncell, cap, lo, thr_dense, qmax;
    // cell test of the candidates: the first (up to 4) checked dimensions of
    // S, packed 16 bits each into one word per sorted position; the others
    // (restmask, positions in S) are tested one by one
    int nchk4;
    int chk[4];
    uint64_t restmask;
    // phi mode (rectangular = false): dbit = 2 (dense = bit 1 of the stored
    // decision bits), thresholds of the weak and the dense test; count mode: dbit = 0
    int dbit;
    float phi_w, phi_d;
    // positions of S whose eps cells restrict the candidates: the dimensions
    // of S the neighbourhood tree walk matches in level order (all of S in
    // the identity order; ScyTreeNode.cu:910-915)
    uint64_t matchmask;
    const uint16_t *cellE;  // [d][n] eps-grid cells by point (the query's own cells are read here)
};

// Device buffers of one batch's grid indexes.
struct BatchIndex {
    int *keys, *rank;     // [B*n]
    int *hist, *off;      // [nhist]
    int *order;           // [B*n] point id at each sorted position
    float *Ys;            // coordinates in sorted order
    uint16_t *Es;         // eps-grid cells in sorted order (gathered only when a walk tests cells one by one)
    uint8_t *qflag;       // [B*n] query flag per (global) sorted position
    int *qlist;           // global sorted positions of the queries, ascending
    int *nsel;            // [1] number of queries
    int *qstart;          // [B+1] first query of each subspace in qlist
    void *select_temp;
    size_t select_bytes;
    int *nb;              // [B*n*kSampleEdges] first neighbours found by the count passes
    uint64_t *sig;        // [B*n] packed cells of the checked dimensions per sorted position
    uint64_t *sigS;       // [B*n] (index space) the same over the shared dimensions only
    const int *reps;      // subspaces that build an index, in index order
    int nrep;
    int2 *fb;             // phi mode: (subspace, point) pairs to recompute in leaf order
    int *nfb;             // [1] their number
    int fbcap;
    uint8_t *dsorted;     // [B*n] dense flag per sorted position
    int *par;             // [B*n] union-find parents (point ids within each subspace)
    int *giant;           // [B]
    void *scan_temp;
    size_t scan_bytes;
    int nhist;            // sum of ncell + 1
};

// Tree whose s-connection paths set the merge flags of its dimensions >= first.
struct FlagTree {
    const int *spx;
    const uint64_t *spm;  // mw words per path
    const int *nsp;  // device count
    int first, flag_off, mw;
};

// Points of a tree that are queries of subspace b.
struct QuerySeg {
    const int *P;
    int len, b;
};

// Partition of one tree by the cell groups of dimension m (k_part_* per pair).
struct PartPair {
    const int *P;
    const int *spx;
    const uint64_t *spm;  // mw words per path
    const int *gid;      // [cells]
    const int *weakcnt;  // counts of S ∪ {m}
    int *cnt;            // [6G]: points, weak-dense points, paths per group; then cursors
    int *outW;
    int *outx;
    uint64_t *outm;
    int nP, nsp, m, G, bp, mw;
};

// Path prune of one surviving tree.
struct PruneChild {
    dmask_t smask;
    uint64_t tmask;
    const int *W;
    const int *spx;
    const uint64_t *spm;  // mw words per path
    uint64_t *table;
    dmask_t *bits;
    int *outx;
    uint64_t *outm;
    int *nout;
    int nW, nsp, bw, mw;
};

struct CopySeg {
    const int *src;
    int *dst;
    int len, pad;
};

// Largest float s with sqrtf(s) <= eps: the distance test fl(sqrt(sum)) <= eps
// of ClusteringCpu.cu:52-81 is then sum <= s_star without a device sqrt.
float sqrt_threshold(float eps);
// exclusive prefix sum of items ints (inscy_scan.hip); temp of scan_temp_bytes(items)
size_t scan_temp_bytes(int items);
hipError_t launch_exclusive_sum(void *temp, size_t temp_bytes, const int *in, int *out, int items, hipStream_t st);
// stable sort of (uint16 key, int value) pairs by the low `bits` (<= 16) key
// bits; two 8-bit passes above 8 bits go through keys_tmp / vals_tmp
size_t sort_temp_bytes(int n);
hipError_t launch_sort_pairs_u16(const uint16_t *keys, uint16_t *keys_tmp, uint16_t *keys_out, const int *vals,
                                 int *vals_tmp, int *vals_out, int n, int bits, void *temp, size_t temp_bytes,
                                 hipStream_t st);
size_t select_temp_bytes(int items);
// positions of the nonzero bytes of flags[0, nb) in order into out, their
// number into *nsel; flags readable up to nb rounded up to 16 bytes, temp of
// select_temp_bytes(nb) bytes
hipError_t launch_flag_compact(const uint8_t *flags, size_t nb, void *temp, size_t temp_bytes, int *out, int *nsel,
                               hipStream_t st);

// Launchers (all asynchronous on `st`; return the hipError_t of the launches).
// discretise: the near-edge marks of point p are nw 64-bit words at nearw[p * nw]
// (nw = 2 is dmask_t[n]; 1 suffices for d <= 64)
hipError_t launch_discretise(const float *X, int n, int d, int cells, float eps, float *Xt, uint8_t *cellC,
                             uint16_t *cellE, uint64_t *nearw, int nw, int *bad, hipStream_t st);
hipError_t launch_hist1(const uint16_t *cellE, int n, int d, int ecells, int *hist, hipStream_t st);
// occupancy of the c-grid per dimension ([d][cells]), the input of the entropy order
hipError_t launch_histc(const uint8_t *cellC, int n, int d, int cells, int *hist, hipStream_t st);
// |N_S(p)| (capped at `cap`) of the points with cnt == -1 and, when mark is
// given, mark[p] == markid (stat: candidates examined, points counted)
hipError_t launch_subspace_counts(const DevData &D, const SubspaceArg &S, const GridArg &G, float s_star, int cap,
                                  const GridBuffers &B, const int *mark, int markid, int *cnt,
                                  unsigned long long *stat, const KernelTimer &tm, hipStream_t st);
// remaining counts, then the dense components of S (comp = smallest member id or -1)
// (bound: complete counts of a subspace of S, or null; lo: the smaller threshold of S)
// dbit != 0: cnt holds phi decision bits and bit dbit is the dense test
hipError_t launch_subspace_components(const DevData &D, const SubspaceArg &S, const GridArg &G, float s_star,
                                      int cap, int lo, int thr_dense, int dbit, const GridBuffers &B, const int *bound,
                                      int *cnt, int *comp, int *csize, unsigned long long *stat,
                                      const KernelTimer &tm, hipStream_t st);
hipError_t launch_mark(const int *P, const int *nP, int nP_max, int *mark, int markid, hipStream_t st);
hipError_t launch_flags_all(const int *spx, const uint64_t *spm, int mw, const int *nsp, int nsp_max,
                            const uint8_t *cellC, int n, int first, int cells, int *flags, hipStream_t st);
hipError_t launch_partition(const int *P, const int *nP, int nP_max, const int *spx, const uint64_t *spm, int mw,
                            const int *nsp, int nsp_max, const uint8_t *cellC_m, const int *gid, int G,
                            const int *weakcnt, int thr, int wbit, int m, int *cnt, int *outP, int *outW, int *outx,
                            uint64_t *outm, hipStream_t st);
hipError_t launch_path_prune(const DevData &D, dmask_t smask, const int *W, const int *nW, int nW_max,
                             const int *spx, const uint64_t *spm, int mw, const int *nsp, int nsp_max, int *stamp,
                             int tid, dmask_t *bits, uint64_t *table, uint64_t tmask, int *outx, uint64_t *outm,
                             int *nout, hipStream_t st);
hipError_t launch_redundancy_seed(const int *W, const int *nW_dev, int nW, int *stamp, int tid, const Entry *E,
                                  const int *nE, dmask_t smask, const int *comp, int *list, int *nlist, int *mn,
                                  float r, int *keep, uint8_t *lab, hipStream_t st);
hipError_t launch_join(const uint8_t *lab, int n, const int *compS, const int *csizeS, int min_size, float r,
                       Entry *E, const int *nE, dmask_t smask, int eS, uint8_t *inresS, int *list, int *nlist,
                       uint8_t *keepf, uint8_t *drop, uint8_t *fresh, hipStream_t st);
hipError_t launch_set_entry(Entry *E, int idx, const Entry &e, int *nE, int count, hipStream_t st);
// canonical labels of r result subspaces: flags and idx hold r*n + 1 ints,
// labels r*n, counts r (clusters per row)
struct EmitRow {
    const int *comp;
    const uint8_t *inres;
};
// width: bytes per label written (4: int32 labels; 1 / 2: label + 1 as
// uint8 / uint16, for rows with at most 255 / 65535 clusters)
hipError_t launch_emit(const EmitRow *rows, int r, int n, int *flags, int *idx, void *labels, int width, int *counts,
                       void *scan_temp, size_t scan_bytes, hipStream_t st);
hipError_t launch_fill_u8(uint8_t *p, uint8_t v, size_t count, hipStream_t st);
hipError_t launch_fill_i32(int *p, int v, size_t count, hipStream_t st);
hipError_t launch_gather(const int *src, const int *idx, int count, int *out, hipStream_t st);
hipError_t launch_iota(int *p, int n, hipStream_t st);
hipError_t launch_root_paths(const dmask_t *nearm, int n, int *outx, uint64_t *outm, int mw, int *nout, hipStream_t st);

// level-synchronous passes (blk_off: per-segment block offsets, nblk = blk_off[nseg])
hipError_t launch_b_flags(const FlagTree *T, const int *blk_off, int nt, int nblk, const uint8_t *cellC, int n,
                          int cells, int *flags, hipStream_t st);
// grid indexes of nsub subspaces: keys, one scan over all grid cells, scatter
// A grid index sort kept for later batches of the run (the same index
// dimensions sort the points the same way): sorted point ids, each point's
// sorted position and the cell offsets, all local to the index; with the
// batch-side place of one index (obase, hoff) and its cell count.
struct SortCopy {
    int *order, *pos, *off;
    long long obase, hoff;
    int ncell;
};

// need_es: gather the sorted cell columns Es (a launch of the batch runs the
// walk for any k, which tests cells one by one).  miss_reps: the indexes to
// sort (positions in subs), placed first: their cells are the histogram's
// first hmiss entries; stores: sorts to keep (filled from the batch after
// sorting); hits: kept sorts copied in (after the stores, so an index repeated
// within the batch reads the sort just kept).  After it X.keys holds every
// point's local sorted position.
// Sorted coordinate columns kept across batches (same index dimensions and
// dimension): per index column (rep position * kmax + t) and per subspace's
// own column, a kept column to read (src) or a slot to keep a gathered one
// (store), either null.
struct ColumnKeep {
    const float *const *src;
    float *const *store;
    const float *const *own_src;
    float *const *own_store;
    int kmax;
};
hipError_t launch_b_index(const DevData &D, const BatchSub *subs, int nsub, const BatchIndex &X, bool need_es,
                          const int *miss_reps, int nmiss, int hmiss, const SortCopy *stores, int nstore,
                          const SortCopy *hits, int nhit, int max_cell, const ColumnKeep &cols, hipStream_t st);
// queries in sorted order: flag the points of the segments, compact per subspace
hipError_t launch_b_queries(const BatchSub *subs, int nsub, int n, const QuerySeg *segs, const int *blk_off,
                            int nseg, int nblk, const BatchIndex &X, hipStream_t st);
// neighbour counts of the listed queries (blk_off over subspaces), or of every
// uncounted point when blk_off is null (grid.y = subspace)
// phi: rectangular = false, the stored values are decision bits (1 weak-dense,
// 2 dense); points near a threshold are recomputed in leaf order (rank_order)
// kfast in 1..7: every subspace of the launch has k = kfast and at most four
// checked dimensions (the count kernel is then compiled for that k); else 0
hipError_t launch_b_count(const BatchSub *subs, int nsub, const int *blk_off, int nblk, int n, int kmax, int kfast,
                          const BatchIndex &X, float s_star, float eps, bool phi, bool phi_all_exact,
                          const DevData &D, const int *rank_order, unsigned long long *stat, const KernelTimer &tm,
                          hipStream_t st);
// phi decision bits of the points of one subspace without a value (cnt ==
// -1; stamped markid when mark is given), from the grid pass of S: decided
// where a threshold lies outside the rounding band (phirel per term), the
// others listed in fb (capacity fbcap, count in nfb) as (0, point)
hipError_t launch_subspace_phi(const DevData &D, const SubspaceArg &S, const GridArg &G, const GridBuffers &B,
                              float s_star, float eps, double phirel, float phi_w, float phi_d, const int *mark,
                              int markid, int *cnt, int2 *fb, int *nfb, int fbcap, hipStream_t st);
// the listed points' phi summed in the neighbourhood tree's leaf order
// (rank_order), stored in sub's cnt (sub: one device BatchSub with S,
// matchmask, phi_w, phi_d and cnt)
hipError_t launch_phi_exact(const DevData &D, const BatchSub *sub, const int2 *fb, const int *nfb, int fbcap,
                            const int *rank_order, float eps, float s_star, hipStream_t st);
// dense components of nsub subspaces whose counts are complete
hipError_t launch_b_components(const BatchSub *subs, int nsub, int n, int kmax, int kfast, const BatchIndex &X,
                               float s_star, hipStream_t st);
// weak test: count >= thr, or bit wbit of the phi decision bits when wbit != 0;
// lists are written only for groups with at least min_size weak-dense points
hipError_t launch_b_partition(const PartPair *pairs, const int *blk_off, int npairs, int nblk, const uint8_t *cellC,
                              int n, int thr, int wbit, int min_size, hipStream_t st);
// points sorted by their eps cells over all dimensions in `level` order, ties
// by id: the leaf order of the reference's neighbourhood tree (d stable passes)
hipError_t launch_rank_order(const uint16_t *cellE, int n, int d, int ecells, const int *level, int *perm, int *tmp,
                             uint16_t *keys, uint16_t *keys2, void *temp, size_t temp_bytes, hipStream_t st);
size_t rank_order_temp_bytes(int n);
// path prune of many trees (blk_w over their points, blk_p over their paths)
hipError_t launch_b_prune(const DevData &D, const PruneChild *C, int nc, const int *blk_w, int nblk_w,
                          const int *blk_p, int nblk_p, hipStream_t st);
hipError_t launch_b_copy(const CopySeg *segs, const int *blk_off, int nseg, int nblk, hipStream_t st);
hipError_t launch_b_entries(Entry *E, int at, const Entry *src, int count, int *nE, hipStream_t st);
// dst[src[i]] = pad for every segment (the tree stamps of the replay)
hipError_t launch_b_stamp(const CopySeg *segs, const int *blk_off, int nseg, int nblk, hipStream_t st);
// replay: pruneRedundancy of one tree (keepv[tree] = decision; rowflag: 2 *
// nlist zeroed words and done: one zeroed word, both left zeroed by the
// launch), join of one (parent, dimension) pair (added: one scratch word)
hipError_t launch_rep_redundancy(const Entry *E, const int *list, int nlist, const int *tstamp, const int *W,
                                 int tree, int nW, float r, int *rowflag, unsigned *done, int *keepv, hipStream_t st);
// One step of the staged replay: pruneRedundancy of `tree` (kind 0 and 1),
// and/or the join J of the kept trees [t_lo, t_hi) into subspace S (kind 0:
// the pair's only surviving child is `tree`; kind 2: a pair with several).
// Steps of one stage touch pairwise incomparable subspaces, so they commute.
// flag: S's count array, reused for the join's per-component flags
// (candidate -(2+3J), dropped -(3+3J), joined -(4+3J)); clist: candidate
// roots (capacity ccap); words: [0] candidates listed, [1] clusters added.
struct RepOp {
    const int *list;
    const int *tstamp;
    const int *W;
    int *flag;
    const int *comp;
    const int *csize;
    uint8_t *inres;
    int *rowflag;
    unsigned *done;
    int *words;
    int *clist;
    int nlist, nW, tree, t_lo, t_hi, entry, J, kind, ccap, pad;
};
// the staged replay's passes over nops steps; each takes blk_off (per-step
// block offsets) over the steps it serves
hipError_t launch_rep_stage(const RepOp *ops, int nops, const int *blk_red, int nblk_red, const int *blk_cand,
                            int nblk_cand, const int *blk_chk, int nblk_chk, const int *blk_mem, int nblk_mem,
                            Entry *E, float r, int min_size, int n, int *keepv, hipStream_t st);
// blocks per step of each pass (the host sizes the offsets with these)
int rep_red_blocks(int nlist, int nW);
int rep_check_blocks(int nlist, int n);
int rep_mem_blocks(int n);

hipError_t launch_fill_u64(uint64_t *p, uint64_t v, size_t count, hipStream_t st);

// ---- learned pruner
constexpr int kMlpIn = 64, kMlpHidden = 128;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden + kMlpHidden + 1;
constexpr int kHistBins = 16;  // value histogram of a tree's points in the candidate dimension

// One (tree, dimension m) pair scored before its neighbour counts: the
// tree's points W, the complete counts of its subspace S (null for the
// root) and the cell -> merge group map of dimension m.
struct FeatPair {
    const int *W;
    const int *cntS;  // counts of S, capped at cap
    const int *gid;   // [cells]
    dmask_t mask;     // S | {m}
    int nW, pre, m, k, first, ngroups, bp, cap;
};
// acc: per pair kHistBins + cells + 2 ints (value histogram, cell histogram,
// points with |N_S| >= cap and >= (weak_thr + cap) / 2); accs: per pair the
// sum of |N_S| (u64).  feat: [npairs][kMlpIn].
hipError_t launch_pair_features(const FeatPair *pairs, const int *blk_off, int npairs, int nblk, const float *Xt,
                                const uint8_t *cellC, int n, int d, int cells, int min_size, int weak_thr, int *acc,
                                unsigned long long *accs, float *feat, hipStream_t st);
// network (bf16 parameters P, kMlpParams, layout W1 b1 W2 b2 w3 b3)
hipError_t launch_mlp_predict(const float *X, long long N, const uint16_t *P, float *prob, hipStream_t st);
// gradient of the weighted mean binary cross-entropy: part holds nblocks rows
// of kMlpParams + 1 floats, grad receives kMlpParams + 1 (the last the loss)
hipError_t launch_mlp_grad(const float *X, const float *Y, const float *WT, long long N, float inv_wsum,
                           const uint16_t *P, float *part, int nblocks, float *grad, hipStream_t st);
hipError_t launch_mlp_adam(float *master, const float *grad, float *m, float *v, uint16_t *P, int step, float lr,
                           hipStream_t st);
hipError_t launch_mlp_round(const float *master, uint16_t *P, hipStream_t st);

// hit[t] |= 1 when a point of tree t's list lies in a kept cluster of its subspace
struct TreeHit {
    const int *W;
    const int *comp;
    const uint8_t *inres;
    int nW, out;
};
hipError_t launch_tree_hits(const TreeHit *trees, const int *blk_off, int ntrees, int nblk, int *hit, hipStream_t st);

constexpr int kRangeWords = 31;      // per query in LDS: 9 range starts, 10 prefix sums, count, own run, 10 group sums
constexpr int kBlockThreads = 256;   // threads per block of every pass
constexpr int kQueriesPerBlock = 64;  // queries per block of the batched neighbour passes
constexpr int kSampleEdges = 2;       // neighbours recorded per point by the count walk (Afforest's sample rounds)
constexpr int kPartPerThread = 8;     // elements per thread of the batched partition passes

// ------------------------------------------------------------ allocation-free neighbour counts
// One file-scope device array per device (inscy_noalloc.hip) holds every
// buffer of the pass: bytes needed for n points, k dimensions and ecells
// cells per dimension; the array's address on the current device (looked up
// once per device); the pass itself on `st` (counts of the query ids,
// -1 for an id outside [0, n)).
constexpr size_t kNoAllocScratchBytes = (size_t) 64 << 20;
size_t noalloc_scratch_need(int n, int k, int ecells);
hipError_t noalloc_scratch(unsigned char **out);
hipError_t launch_neighbour_counts_noalloc(unsigned char *base, const float *X, int n, int d, float eps,
                                           float s_star, const SubspaceArg &S, const int *queries, int nq, int *out,
                                           hipStream_t st);

}  // namespace inscy
