// Host driver and C ABI of the MI355X INSCY engine.
//
// The reference drives INSCY from the host with one pointer tree per
// recursion node (INSCYCPU3Weak, src/algorithms/inscy/InscyNodeCpu.cu:159-235)
// or with device SCY-tree arrays plus many synchronous scalar reads per
// restricted tree (InscyArrayGpu4/5, InscyArrayGpuMulti2.cu:569-836).  Here a
// tree lives on the device as three flat lists:
//   P   — its points (int32),
//   SPx/SPm — its s-connection paths: origin point and a bitmask of the
//         dimensions at which that point's chain branches (only dimensions
//         above the last restricted one are kept, the only ones a descendant
//         can still restrict),
// plus the host-side bitmask S of its restricted dimensions.  restrict +
// mergeWithNeighbors become one flag pass and one filter pass, the weak-density
// prune is a neighbour-count pass plus compaction, and clustering works on the
// dense components of S over the whole data set, computed once per subspace.
// The host keeps the depth-first order of the reference because
// pruneRedundancy and join read the partial result.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include <sys/mman.h>

#include "../../include/inscy.h"
#include "inscy_internal.h"
#include "inscy_memory.h"

// Host array for the result labels: 2 MiB aligned and advised for huge pages,
// so that first touch of a multi-GB result costs few page faults.
struct HostFree {
    void operator()(int32_t *p) const { std::free(p); }
};
inline int32_t *alloc_labels(size_t count) {
    const size_t align = (size_t) 2 << 20;
    const size_t bytes = ((std::max<size_t>(count, 1) * sizeof(int32_t)) + align - 1) & ~(align - 1);
    void *p = nullptr;
    if (posix_memalign(&p, align, bytes) != 0) throw std::bad_alloc();
    (void) madvise(p, bytes, MADV_HUGEPAGE);
    return static_cast<int32_t *>(p);
}

struct inscy_result {
    int64_t n = 0;
    std::vector<std::vector<int32_t>> subspaces;
    std::unique_ptr<int32_t[], HostFree> labels;  // [subspaces][n], canonical
    inscy_stats stats{};
    // INSCY_MODE_COLLECT: the learned pruner's training pairs
    std::vector<float> pair_feat;      // [pairs][kMlpIn]
    std::vector<uint8_t> pair_label;   // [pairs]
    std::vector<int32_t> pair_info;    // [pairs][4]: |S| of the tree, m, tree points, tree index
    std::vector<int64_t> pair_cost;    // [pairs][2]: own points counted, with every pair below
};

namespace inscy {

struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            throw Failure(e_ == hipErrorOutOfMemory ? INSCY_ERR_MEMORY : INSCY_ERR_DEVICE,                 \
                          std::string(#expr) + ": " + hipGetErrorString(e_));                              \
    } while (0)

// ------------------------------------------------------------ thresholds
// ClusteringCpu.cu:27-38,136-183 with PI = 3.14; float results of
// double-promoted pow() as in the reference.
inline float gamma_half(int m) {
    if (m == 2) return 1.0f;
    if (m == 1) return (float) std::sqrt(3.14);
    return (float) ((m / 2.0 - 1.0) * (double) gamma_half(m - 2));
}
inline float ball(int k) {
    float v = (float) std::pow(3.14, k / 2.0);
    return v / gamma_half(k + 2);
}
inline float expden(int k, float eps, int n) {
    float nc = (float) n * ball(k);
    float v = (float) ((double) nc * std::pow((double) eps, (double) k));
    return (float) ((double) v / std::pow(1.0, (double) k));
}
// alpha(k), omega(k) (ClusteringCpu.cu:36-38,152-157): the phi thresholds of
// the non-rectangular density, float results of double-promoted pow()
inline float alpha_thr(int k, float eps, int n) {
    float v = (float) ((double) (2 * n) * std::pow((double) eps, (double) k) * (double) ball(k));
    return (float) ((double) v / (std::pow(1.0, (double) k) * (double) (k + 2)));
}
inline float omega_thr(int k) { return (float) (2.0 / (k + 2.0)); }

// Smallest integer count c with (float)c >= t (counts are compared as floats
// in the reference, ScyTreeNode.cu:616 and ClusteringCpu.cu:176-183).
inline int count_threshold(float t) {
    if (!(t > 0.0f)) return 0;
    double c = std::ceil((double) t);
    if (c > 2147483647.0) return INT_MAX;
    int v = (int) c;
    while (v > 0 && (float) (v - 1) >= t) --v;
    while ((float) v < t) ++v;
    return v;
}

// ------------------------------------------------------------ memory
// Stack-disciplined bump allocator for the per-tree lists of the depth-first
// recursion (a child's lists are released when its subtree is done).
struct Arena {
    char *base = nullptr;
    size_t cap = 0, top = 0;
    ~Arena() {
        if (base) dev_free(base);
    }
    void init(size_t bytes) {
        cap = bytes;
        HIP_OK(dev_alloc(&base, bytes));
    }
    template <class T>
    T *take(size_t count) {
        size_t at = (top + 255) & ~(size_t) 255;
        size_t bytes = std::max<size_t>(count * sizeof(T), 1);
        if (at + bytes > cap) throw Failure(INSCY_ERR_MEMORY, "device arena exhausted");
        top = at + bytes;
        return reinterpret_cast<T *>(base + at);
    }
    size_t mark() const { return top; }
    void release(size_t m) { top = m; }
};

// Grow-only slabs for per-subspace arrays kept for the whole run.
struct Slabs {
    std::vector<void *> chunks;
    char *cur = nullptr;
    size_t left = 0, chunk = 256u << 20;
    ~Slabs() {
        for (void *p : chunks) dev_free(p);
    }
    template <class T>
    T *take(size_t count) {
        size_t bytes = ((count * sizeof(T)) + 255) & ~(size_t) 255;
        if (bytes > left) {
            size_t sz = std::max(chunk, bytes);
            void *p = nullptr;
            HIP_OK(dev_alloc(&p, sz));
            chunks.push_back(p);
            cur = static_cast<char *>(p);
            left = sz;
        }
        T *r = reinterpret_cast<T *>(cur);
        cur += bytes;
        left -= bytes;
        return r;
    }
};

// Bump allocator over device chunks that are kept across reset(): the
// per-level and per-batch buffers of the level-synchronous driver.  Reuse
// after reset() is safe because every user of the memory runs on one stream.
struct Pool {
    struct Chunk {
        char *p;
        size_t cap;
    };
    std::vector<Chunk> chunks;
    size_t ci = 0, top = 0, chunk = 64u << 20;
    ~Pool() {
        for (auto &c : chunks) dev_free(c.p);
    }
    template <class T>
    T *take(size_t count) {
        size_t bytes = ((std::max<size_t>(count * sizeof(T), 1)) + 255) & ~(size_t) 255;
        while (ci < chunks.size() && top + bytes > chunks[ci].cap) {
            ++ci;
            top = 0;
        }
        if (ci == chunks.size()) {
            size_t sz = std::max(chunk, bytes);
            void *p = nullptr;
            HIP_OK(dev_alloc(&p, sz));
            chunks.push_back({static_cast<char *>(p), sz});
            top = 0;
        }
        T *r = reinterpret_cast<T *>(chunks[ci].p + top);
        top += bytes;
        return r;
    }
    void reset() {
        ci = 0;
        top = 0;
    }
};

// Pinned host memory that grows on demand (contents are not kept).
struct HostBuf {
    char *p = nullptr;
    size_t cap = 0;
    ~HostBuf() {
        if (p) host_free(p);
    }
    char *get(size_t bytes) {
        if (bytes > cap) {
            if (p) host_free(p);
            p = nullptr;
            size_t sz = std::max<size_t>(bytes, (size_t) 1 << 20);
            HIP_OK(host_alloc(&p, sz));
            cap = sz;
        }
        return p;
    }
};

struct TreeRef {
    dmask_t smask;
    const int *P;      // device point ids
    const int *nP;     // device count
    int nP_max;        // host bound (exact where the host knows it)
    const int *spx;    // device path origins
    const uint64_t *spm;  // their masks, mw words each
    const int *nsp;
    int nsp_max;
    const uint64_t *chain = nullptr;  // permuted-order driver: chain words instead of spm
};

// What the algorithm needs about one subspace S, computed on first use:
// neighbour counts of every point (the prune reads them for its region), then,
// once a tree of S survives its prune, the dense components of S over the
// whole data set and its row in the device result table.
struct Subspace {
    int *cnt = nullptr;        // [n] |N_S(p)| capped, -1 = not counted yet
    int *comp = nullptr;       // [n] component root (smallest id) or -1
    int *csize = nullptr;      // [n] component sizes indexed by root
    uint8_t *inres = nullptr;  // [n] component kept in the result of S
    int *tstamp = nullptr;     // [n] surviving tree of S holding the point (level-synchronous driver)
    int entry = -1;            // row in the device result table
};

class Engine {
public:
    Engine(const inscy_params &p, int n, int d, hipStream_t st) : prm(p), n(n), d(d), st(st) {}

    ~Engine() {
        // blocks go back to the shared caches: nothing queued may still use them
        (void) hipStreamSynchronize(st);
        for (auto &t : timers) {
            event_pool().put(t.start);
            event_pool().put(t.stop);
        }
        for (void *p : {(void *) Xt, (void *) cellC, (void *) cellE, (void *) nearm, (void *) small, (void *) stamp,
                        (void *) qmark, (void *) entries, (void *) scratch, (void *) rank_perm,
                        (void *) d_level})
            if (p) dev_free(p);
        if (pin) host_free(pin);
    }

    // ---------------------------------------------------------- set-up
    void setup(const float *d_X) {
        cells = prm.number_of_cells;
        ecells = (int) std::ceil(1.0 / (double) prm.neighborhood_size);
        if (ecells > 65535) throw Failure(INSCY_ERR_UNSUPPORTED, "neighborhood_size below 1/65535 is not supported");
        s_star = sqrt_threshold(prm.neighborhood_size);
        weak_thr = count_threshold(std::max(prm.F * expden(d, prm.neighborhood_size, n), (float) prm.num_obj));
        for (int k = 0; k <= d; ++k)
            dense_thr.push_back(
                count_threshold(std::max(prm.F * expden(k, prm.neighborhood_size, n), (float) prm.num_obj)));
        // rectangular = false: phi against max(F alpha, num_obj omega), weak
        // with all d dimensions, dense with |S| (ClusteringCpu.cu:166-176)
        phi_mode = prm.rectangular == 0;
        phi_weak = std::max(prm.F * alpha_thr(d, prm.neighborhood_size, n), (float) prm.num_obj * omega_thr(d));
        for (int k = 0; k <= d; ++k)
            phi_dense.push_back(
                std::max(prm.F * alpha_thr(k, prm.neighborhood_size, n), (float) prm.num_obj * omega_thr(k)));

        mw = d <= 64 ? 1 : 2;
        size_t nd = (size_t) n * d;
        HIP_OK(dev_alloc(&Xt, nd * sizeof(float)));
        HIP_OK(dev_alloc(&cellC, nd));
        HIP_OK(dev_alloc(&cellE, nd * sizeof(uint16_t)));
        HIP_OK(dev_alloc(&nearm, (size_t) n * sizeof(dmask_t)));
        HIP_OK(dev_alloc(&small, kSmallWords * sizeof(unsigned long long)));
        HIP_OK(dev_alloc(&stamp, (size_t) n * sizeof(int)));
        HIP_OK(dev_alloc(&qmark, (size_t) n * sizeof(int)));
        HIP_OK(host_alloc(&pin, pin_bytes()));
        grow_entries(1 << 12);
        // grid index size: C^j cells with j <= 3 and C^j <= max(8n, 4096)
        grid_cap = std::max<long long>(8LL * n, 4096);
        grid_cap = std::max<long long>(std::min<long long>(grid_cap, 1LL << 24), ecells);
        scan_bytes = scan_temp_bytes((int) grid_cap + 1);
        // arena: per recursion level the partitions of all children of one tree
        // (points, weak-dense points, paths) plus one child's pruned paths,
        // and the scratch of one subspace grid pass
        size_t per_level = (size_t) 24 * d * n + (size_t) 64 * n;
        size_t pass = (size_t) n * (32 + (size_t) d * 6) + (size_t) (grid_cap + 1) * 8 + scan_bytes + (1u << 20);
        size_t want = per_level * (size_t) (d + 2) + table_bytes_max() + pass + (size_t) n * 16 + (64u << 20);
        // the level-synchronous driver keeps its lists in pools; its arena holds
        // the replay's label flags (one per open dimension) and the emit buffers
        if (levels_mode) want = (size_t) n * (d + 8) + pass + (64u << 20);
        if (const size_t free_b = device_available_bytes() / (size_t) mem_share) want = std::min(want, free_b / 3);
        arena.init(want);

        int *hist = arena.take<int>((size_t) d * ecells);
        HIP_OK(launch_fill_i32(reinterpret_cast<int *>(small), 0, 2 * kSmallWords, st));
        HIP_OK(launch_discretise(d_X, n, d, cells, prm.neighborhood_size, Xt, cellC, cellE,
                                 reinterpret_cast<uint64_t *>(nearm), 2, reinterpret_cast<int *>(small), st));
        HIP_OK(launch_hist1(cellE, n, d, ecells, hist, st));
        // c-grid occupancy: the input of the entropy order only
        int *histc = arena.take<int>((size_t) d * cells);
        if (prm.entropy_order != 0) HIP_OK(launch_histc(cellC, n, d, cells, histc, st));
        HIP_OK(launch_fill_i32(stamp, -1, n, st));
        HIP_OK(launch_fill_i32(qmark, -1, n, st));
        std::vector<int> h((size_t) d * ecells), hc((size_t) d * cells);
        HIP_OK(hipMemcpyAsync(pin, small, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(h.data(), hist, h.size() * sizeof(int), hipMemcpyDeviceToHost, st));
        if (prm.entropy_order != 0)
            HIP_OK(hipMemcpyAsync(hc.data(), histc, hc.size() * sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        if (reinterpret_cast<int *>(pin)[0] != 0)
            throw Failure(INSCY_ERR_ARGUMENT, "X must be finite with every value in [0, 1] (normalize first)");
        // selectivity of each dimension as an index: expected candidates of a
        // query in the three-cell slab, sum_c h_c (h_{c-1} + h_c + h_{c+1})
        score.assign(d, 0.0);
        for (int k = 0; k < d; ++k) {
            const int *hk = h.data() + (size_t) k * ecells;
            for (int c = 0; c < ecells; ++c) {
                double nb = hk[c] + (c > 0 ? hk[c - 1] : 0) + (c + 1 < ecells ? hk[c + 1] : 0);
                score[k] += (double) hk[c] * nb;
            }
        }
        // SCY-tree level order (get_subspace_order, inscy_map.cpp:61-121):
        // identity, or dimensions by the entropy of their c-grid occupancy
        // (float p*log(p) sums, so an empty cell gives NaN as in the reference),
        // ascending, reversed for entropy_order = -1
        level.resize(d);
        for (int k = 0; k < d; ++k) level[k] = k;
        if (prm.entropy_order != 0) {
            std::vector<std::pair<float, int>> ent(d);
            for (int k = 0; k < d; ++k) {
                float e = 0.0f;
                for (int c = 0; c < cells; ++c) {
                    const float p = (float) hc[(size_t) k * cells + c] / (float) n;
                    e -= p * std::log(p);
                }
                ent[k] = {e, k};
            }
            std::sort(ent.begin(), ent.end());
            for (int k = 0; k < d; ++k) level[k] = ent[k].second;
            if (prm.entropy_order == -1) std::reverse(level.begin(), level.end());
        }
        lvl_of.assign(d, 0);
        for (int t = 0; t < d; ++t) lvl_of[level[t]] = t;
        HIP_OK(dev_alloc(&d_level, (size_t) 2 * d * sizeof(int)));
        std::vector<int> both(level);
        both.insert(both.end(), lvl_of.begin(), lvl_of.end());
        HIP_OK(hipMemcpyAsync(d_level, both.data(), both.size() * sizeof(int), hipMemcpyHostToDevice, st));
        sync();  // `both` is pageable and goes out of scope
        if (phi_mode) {
            // leaf order of the neighbourhood tree, whose levels follow the same order
            HIP_OK(dev_alloc(&rank_perm, (size_t) n * sizeof(int)));
            const size_t tb = rank_order_temp_bytes(n);
            int *tmp = arena.take<int>(n);
            uint16_t *k1 = arena.take<uint16_t>(n);
            uint16_t *k2 = arena.take<uint16_t>(n);
            void *temp = arena.take<char>(tb);
            HIP_OK(launch_rank_order(cellE, n, d, ecells, level.data(), rank_perm, tmp, k1, k2, temp, tb, st));
            sync();
        }
        arena.release(0);
        D = DevData{Xt, cellC, cellE, d_level, d_level + d, n, d, cells, ecells};
    }

    // ---------------------------------------------------------- run
    // level-synchronous driver (INSCY_MODE_BATCHED), same result
    void run_levels(inscy_result &out);
    bool levels_mode = false;
    // permuted SCY-tree level order (entropy_order != 0): the depth-first
    // driver over s-connection chains (inscy_perm.hip)
    void run_perm(inscy_result &out);
    // learned pruner (INSCY_MODE_LEARNED): device bf16 parameters, the score
    // below which a (tree, dimension) pair is skipped, and the smallest |S| of
    // a tree whose pairs may be skipped; pair collection (INSCY_MODE_COLLECT)
    const uint16_t *pruner = nullptr;
    float prune_thr = 0.f;
    int prune_min_depth = 1;
    bool collect = false;
    bool phi_all_exact = false;  // test hook: every phi by the leaf-order recomputation
    bool generic_walk = false;   // test hook: no neighbour walk compiled for its k
    long long keep_budget_override = -1;  // test hook: bytes kept across batches (< 0: the batch budget)
    int mem_share = 1;  // engines running at once on the device (each budgets 1/mem_share of the free bytes)
    int max_batch_override = 0;  // inscy_set_tuning: child subspaces per batch (0: the default)

    void run(inscy_result &out) {
        int *P = arena.take<int>(n);
        int *nP = arena.take<int>(1);
        int *spx = arena.take<int>(n);
        uint64_t *spm = arena.take<uint64_t>((size_t) n * mw);
        int *nsp = arena.take<int>(1);
        HIP_OK(launch_iota(P, n, st));
        HIP_OK(launch_fill_i32(nP, n, 1, st));
        HIP_OK(launch_root_paths(nearm, n, spx, spm, mw, nsp, st));
        TreeRef root{(dmask_t) 0, P, nP, n, spx, spm, nsp, n};
        auto t0 = std::chrono::steady_clock::now();
        recurse(root, 0);
        sync();
        auto t1 = std::chrono::steady_clock::now();
        emit(out);
        stats.expand_seconds = std::chrono::duration<double>(t1 - t0).count();
        stats.emit_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    }

    // dense components of one subspace (kernel-level entry point)
    const Subspace &subspace(dmask_t smask) { return ensure_components(smask); }

    // exact |N_S(p)| of every point, without the threshold cap (kernel-level entry point)
    const int *exact_counts(dmask_t smask) {
        Subspace &s = sub_of(smask);
        SubspaceArg S = arg_of(smask);
        GridArg G = grid_of(S);
        size_t mk = arena.mark();
        GridBuffers B = grid_buffers(S, G);
        HIP_OK(launch_subspace_counts(D, S, G, s_star, INT_MAX, B, nullptr, 0, s.cnt, nullptr, KernelTimer{}, st));
        arena.release(mk);
        return s.cnt;
    }

    inscy_stats stats{};

    struct TraceRow {
        long long mask;
        int lo, hi, pre, post, pass, redundant;
    };
    bool trace_on = false;
    bool profile_on = false;
    std::vector<TraceRow> trace;

private:
    // ---------------------------------------------------------- helpers
    // `small` device words: [0] input check flag, [16..24) path-prune scratch,
    // [kStatAt..) 64 slot pairs (16 words apart) of neighbour-count statistics
    static constexpr int kStatAt = 32;
    static constexpr int kSmallWords = kStatAt + 16 * 64;
    size_t pin_bytes() const {
        return std::max<size_t>((size_t) n * 8 + (size_t) 2 * d * 256 * 8 + (1u << 16), 1u << 20);
    }
    static size_t table_slots(int nW, int bits) {
        size_t want = (size_t) std::max(1, nW) * (size_t) std::max(1, bits) * 2;
        size_t s = 1024;
        while (s < want) s <<= 1;
        return s;
    }
    size_t table_bytes_max() const { return table_slots(n, d) * sizeof(uint64_t); }
    int *pin_i(size_t off = 0) { return reinterpret_cast<int *>(pin) + off; }

    void sync() {
        flush_uploads();
        HIP_OK(hipStreamSynchronize(st));
        up_top = 0;  // every staged upload has been consumed
    }
    // uploads between defer_uploads() and flush_uploads() whose staging and
    // device places are contiguous (consecutive takes of one pool) go to the
    // device in one copy; flush before the launch that reads them
    void defer_uploads() { up_defer = true; }
    void flush_uploads() {
        if (pend_bytes) HIP_OK(hipMemcpyAsync(pend_dst, up.p + pend_src, pend_bytes, hipMemcpyHostToDevice, st));
        pend_bytes = 0;
        up_defer = false;
    }

    // next event pair when profiling, else an empty timer
    KernelTimer timer() {
        if (!profile_on) return KernelTimer{};
        KernelTimer t;
        HIP_OK(event_pool().get(&t.start));
        HIP_OK(event_pool().get(&t.stop));
        timers.push_back(t);
        return t;
    }

    // Device result table and the scratch of its kernels, grown on demand.
    void grow_entries(int want) {
        if (want <= ecap) return;
        int cap = std::max(want, 2 * ecap);
        Entry *ne = nullptr;
        int *ns = nullptr;
        HIP_OK(dev_alloc(&ne, (size_t) cap * sizeof(Entry)));
        // scratch: nE, nlist, keep, pad | list[cap] | mn[cap] | keepf[n], drop[n], fresh[n] (bytes)
        HIP_OK(dev_alloc(&ns, (size_t) (8 + 2 * cap) * sizeof(int) + (size_t) 3 * n + 256));
        if (entries) {
            sync();
            HIP_OK(hipMemcpy(ne, entries, (size_t) nentries * sizeof(Entry), hipMemcpyDeviceToDevice));
            dev_free(entries);
            dev_free(scratch);
        }
        entries = ne;
        scratch = ns;
        ecap = cap;
        d_nE = scratch;
        d_nlist = scratch + 1;
        d_keep = scratch + 2;
        d_list = scratch + 8;
        d_mn = scratch + 8 + cap;
        keepf = reinterpret_cast<uint8_t *>(scratch + 8 + 2 * cap);
        dropf = keepf + n;
        freshf = dropf + n;
        HIP_OK(launch_fill_i32(d_mn, INT_MAX, cap, st));
        HIP_OK(launch_fill_i32(d_nE, nentries, 1, st));
    }

    SubspaceArg arg_of(dmask_t smask) const {
        if (dcount(smask) > kMaxSubDims)
            throw Failure(INSCY_ERR_UNSUPPORTED, "a subspace of more than 64 dimensions is not supported");
        SubspaceArg S{};
        S.k = 0;
        for (int k = 0; k < d; ++k)
            if (dhas(smask, k)) S.dims[S.k++] = k;
        return S;
    }

    // positions of S that the neighbourhood tree walk matches: the longest
    // prefix of S (ascending dimension ids) whose levels increase
    // (ScyTreeNode.cu:910-915); every other dimension of S is tested by the
    // distance alone
    uint64_t match_of(const SubspaceArg &S) const {
        uint64_t m = 0;
        int si = 0;
        for (int t = 0; t < d && si < S.k; ++t)
            if (level[t] == S.dims[si]) m |= 1ull << si++;
        return m;
    }

    // grid index of S: the (up to 3) most selective matched dimensions, not
    // counting position `exclude` (the child's own dimension of a shared index);
    // fine (the level-synchronous driver): an index of at most two dimensions
    // sorts its last one in kFineBuckets buckets per cell, so that a query's
    // runs can leave out the buckets farther than eps (its cells stay few)
    GridArg grid_of(const SubspaceArg &S, int exclude = -1, bool fine = false) const {
        GridArg G{};
        const uint64_t matched = match_of(S);
        std::vector<int> ts;
        for (int t = 0; t < S.k; ++t)
            if (((matched >> t) & 1ull) && t != exclude) ts.push_back(t);
        std::stable_sort(ts.begin(), ts.end(), [&](int a, int b) { return score[S.dims[a]] < score[S.dims[b]]; });
        G.C = ecells;
        G.j = 1;
        long long cellsj = ecells;
        while (G.j < 3 && G.j < (int) ts.size() && cellsj * ecells <= grid_cap) {
            cellsj *= ecells;
            G.j++;
        }
        G.F = fine && G.j <= 2 && cellsj * kFineBuckets <= grid_cap ? kFineBuckets : 1;
        G.checkmask = matched;
        for (int a = 0; a < G.j; ++a) {
            G.tI[a] = ts[a];
            G.checkmask &= ~(1ull << ts[a]);
        }
        return G;
    }

    GridBuffers grid_buffers(const SubspaceArg &S, const GridArg &G) {
        long long ncell = 1;
        for (int a = 0; a < G.j; ++a) ncell *= G.C;
        GridBuffers B{};
        B.keys = arena.take<int>(n);
        B.rank = arena.take<int>(n);
        B.hist = arena.take<int>(ncell + 1);
        B.off = arena.take<int>(ncell + 1);
        B.order = arena.take<int>(n);
        B.Ys = arena.take<float>((size_t) S.k * n);
        B.Es = arena.take<uint16_t>((size_t) S.k * n);
        B.dsorted = arena.take<uint8_t>(n);
        B.par = arena.take<int>(n);
        B.giant = arena.take<int>(1);
        B.scan_temp = arena.take<char>(scan_bytes);
        B.scan_bytes = scan_bytes;
        return B;
    }

    // counts are compared only with the weak threshold (prune) and the dense
    // threshold of |S| (clustering), so they are needed up to the larger one
    int cap_of(int k) const { return std::max(std::max(weak_thr, dense_thr[k]), 1); }

    Subspace &sub_of(dmask_t smask) {
        Subspace &s = subs[smask];
        if (!s.cnt) {
            s.cnt = slabs.take<int>(n);
            HIP_OK(launch_fill_i32(s.cnt, -1, n, st));
        }
        return s;
    }

    // phi mode (rectangular = false): the decision bits of the points of S
    // without a value, stamped markid when mark is given (the permuted-order
    // driver's density): the grid pass decides a point whose thresholds lie
    // outside the rounding band of its double scan-order sum (2^-22 per term,
    // as in the level-synchronous passes); the others are summed again in the
    // neighbourhood tree's leaf order
    void phi_points(dmask_t smask, Subspace &s, const int *mark, int markid) {
        BatchSub x;
        std::memset(&x, 0, sizeof(x));
        x.S = arg_of(smask);
        x.matchmask = match_of(x.S);
        x.phi_w = phi_weak;
        x.phi_d = phi_dense[x.S.k];
        x.dbit = 2;
        x.cnt = s.cnt;
        const GridArg G = grid_of(x.S);
        size_t mk = arena.mark();
        BatchSub *dx = arena.take<BatchSub>(1);
        put(dx, &x, sizeof(x));
        int2 *fb = arena.take<int2>(n);
        int *nfb = arena.take<int>(1);
        GridBuffers B = grid_buffers(x.S, G);
        flush_uploads();
        const double phirel = phi_all_exact ? 1e30 : 1.0 / 4194304.0;
        HIP_OK(launch_subspace_phi(D, x.S, G, B, s_star, prm.neighborhood_size, phirel, x.phi_w, x.phi_d, mark, markid,
                                   s.cnt, fb, nfb, n, st));
        HIP_OK(launch_phi_exact(D, dx, fb, nfb, n, rank_perm, prm.neighborhood_size, s_star, st));
        arena.release(mk);  // stream order keeps the buffers for the queued kernels
    }

    // |N_S(p)| of the points stamped with markid that are not counted yet
    // (phi mode: their decision bits).
    Subspace &ensure_counts(dmask_t smask, int markid) {
        Subspace &s = sub_of(smask);
        if (s.entry >= 0) return s;  // all points counted by the components pass
        if (phi_mode) {
            phi_points(smask, s, qmark, markid);
            return s;
        }
        SubspaceArg S = arg_of(smask);
        GridArg G = grid_of(S);
        size_t mk = arena.mark();
        GridBuffers B = grid_buffers(S, G);
        HIP_OK(launch_subspace_counts(D, S, G, s_star, cap_of(S.k), B, qmark, markid, s.cnt, small + kStatAt, timer(),
                                      st));
        arena.release(mk);
        return s;
    }

    // Dense components of S (counting every point not yet counted) and its
    // result-table row, once per subspace.
    Subspace &ensure_components(dmask_t smask) {
        Subspace &s = sub_of(smask);
        if (s.entry >= 0) return s;
        SubspaceArg S = arg_of(smask);
        GridArg G = grid_of(S);
        s.comp = slabs.take<int>(n);
        s.csize = slabs.take<int>(n);
        s.inres = slabs.take<uint8_t>(n);
        int *mem = slabs.take<int>(n);
        // a subspace S' = S minus one dimension whose counts are complete bounds
        // the counts of S from above (exact where below its own cap); phi
        // mode decides every remaining point in leaf order first
        const int *bound = nullptr;
        int lo = std::min(weak_thr, dense_thr[S.k]);
        if (phi_mode) phi_points(smask, s, nullptr, 0);
        for (int t = 0; t < S.k && !bound && !phi_mode; ++t) {
            const dmask_t sub = smask & ~dbit(S.dims[t]);
            auto it = subs.find(sub);
            if (sub && it != subs.end() && it->second.entry >= 0) {
                bound = it->second.cnt;
                lo = std::min(lo, cap_of(S.k - 1));
            }
        }
        size_t mk = arena.mark();
        GridBuffers B = grid_buffers(S, G);
        HIP_OK(launch_subspace_components(D, S, G, s_star, cap_of(S.k), lo, dense_thr[S.k], phi_mode ? 2 : 0, B, bound,
                                          s.cnt, s.comp, s.csize, small + kStatAt, timer(), st));
        arena.release(mk);
        HIP_OK(launch_fill_u8(s.inres, 0, n, st));
        grow_entries(nentries + 1);
        Entry e{smask, s.comp, s.csize, s.inres, mem, 0, 0};
        s.entry = nentries++;
        HIP_OK(launch_set_entry(entries, s.entry, e, d_nE, nentries, st));
        stats.components++;
        return s;
    }

    // ---------------------------------------------------------- recursion
    struct Group {
        int lo, hi;
    };

    void recurse_perm(const TreeRef &T, int first);

    void recurse(const TreeRef &T, int first) {
        stats.calls++;
        if (first >= d) return;
        const int nd = d - first;
        const size_t lvl_mark = arena.mark();

        // restrict + mergeWithNeighbors for every dimension m >= first: the
        // s-connected cells, then the cell groups (ScyTreeNode.cu:837-851)
        int *flags = arena.take<int>((size_t) nd * cells);
        HIP_OK(launch_fill_i32(flags, 0, (size_t) nd * cells, st));
        HIP_OK(launch_flags_all(T.spx, T.spm, mw, T.nsp, T.nsp_max, cellC, n, first, cells, flags, st));
        // the subspaces of the children are counted on this tree's points while
        // the flags travel
        const int markid = ++mark_id;
        HIP_OK(launch_mark(T.P, T.nP, T.nP_max, qmark, markid, st));
        for (int m = first; m < d; ++m) ensure_counts(T.smask | dbit(m), markid);
        HIP_OK(hipMemcpyAsync(pin, flags, (size_t) nd * cells * sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        std::vector<std::vector<Group>> groups(nd);
        std::vector<int> gid((size_t) nd * cells);
        for (int m = first; m < d; ++m) {
            const int *f = pin_i((size_t) (m - first) * cells);
            auto &gs = groups[m - first];
            for (int c = 0; c < cells; ++c) {
                int lo = c;
                while (f[c] && c < cells - 1) ++c;
                for (int x = lo; x <= c; ++x) gid[(size_t) (m - first) * cells + x] = (int) gs.size();
                gs.push_back({lo, c});
            }
        }
        int *d_gid = arena.take<int>(gid.size());
        std::copy(gid.begin(), gid.end(), pin_i());
        HIP_OK(hipMemcpyAsync(d_gid, pin, gid.size() * sizeof(int), hipMemcpyHostToDevice, st));

        // partition points, weak-dense points and paths by group, all dimensions
        struct Part {
            int *cnt, *P, *W, *x;
            uint64_t *mk;
        };
        std::vector<Part> parts(nd);
        for (int m = first; m < d; ++m) {
            const int G = (int) groups[m - first].size();
            Part &pt = parts[m - first];
            pt.cnt = arena.take<int>((size_t) 6 * G);
            pt.P = arena.take<int>(std::max(T.nP_max, 1));
            pt.W = arena.take<int>(std::max(T.nP_max, 1));
            pt.x = arena.take<int>(std::max(T.nsp_max, 1));
            pt.mk = arena.take<uint64_t>((size_t) std::max(T.nsp_max, 1) * mw);
            const Subspace &sub = subs[T.smask | dbit(m)];
            HIP_OK(launch_partition(T.P, T.nP, T.nP_max, T.spx, T.spm, mw, T.nsp, T.nsp_max, cellC + (size_t) m * n,
                                    d_gid + (size_t) (m - first) * cells, G, sub.cnt, weak_thr, 0, m, pt.cnt, pt.P,
                                    pt.W, pt.x, pt.mk, st));
        }
        std::vector<size_t> cnt_at(nd);
        size_t tot = 0;
        for (int m = first; m < d; ++m) {
            cnt_at[m - first] = tot;
            tot += 3 * groups[m - first].size();
        }
        int *hostc = pin_i((size_t) nd * cells + 64);
        for (int m = first; m < d; ++m)
            HIP_OK(hipMemcpyAsync(hostc + cnt_at[m - first], parts[m - first].cnt,
                                  3 * groups[m - first].size() * sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        std::vector<int> counts(hostc, hostc + tot);

        for (int m = first; m < d; ++m) {
            const dmask_t smask = T.smask | dbit(m);
            const auto &gs = groups[m - first];
            const int G = (int) gs.size();
            const int *c3 = counts.data() + cnt_at[m - first];
            const Part &pt = parts[m - first];
            const size_t dim_mark = arena.mark();
            uint8_t *lab = nullptr;
            int offW = 0, offS = 0;
            for (int g = 0; g < G; ++g) {
                const int nPr = c3[g], nW = c3[G + g], nSP = c3[2 * G + g];
                const int *W = pt.W + offW;
                const int *spx = pt.x + offS;
                const uint64_t *spm = pt.mk + (size_t) offS * mw;
                offW += nW;
                offS += nSP;
                stats.trees++;
                const bool pass = nW >= prm.min_size;
                size_t at = trace.size();
                if (trace_on) trace.push_back({(long long) dlo(smask), gs[g].lo, gs[g].hi, nPr, nW, pass ? 1 : 0, -1});
                if (!pass) continue;
                const size_t child_mark = arena.mark();
                const int tid = ++tree_id;
                // s-connection paths surviving the prune
                int *spx2 = arena.take<int>(std::max(nSP, 1));
                uint64_t *spm2 = arena.take<uint64_t>((size_t) std::max(nSP, 1) * mw);
                int *nsp2 = arena.take<int>(1);
                {
                    size_t tmark = arena.mark();
                    size_t slots = table_slots(nW, d - 1 - m);
                    uint64_t *table = arena.take<uint64_t>(nSP > 0 ? slots : 1);
                    HIP_OK(launch_path_prune(D, smask, W, pt.cnt + G + g, nW, spx, spm, mw, pt.cnt + 2 * G + g, nSP,
                                             stamp, tid, reinterpret_cast<dmask_t *>(small + 16), table,
                                             (uint64_t) (slots - 1), spx2, spm2, nsp2, st));
                    arena.release(tmark);  // stream order keeps the table alive for the queued kernels
                }
                // components of S now: the children's counts are bounded by them
                Subspace &sub = ensure_components(smask);
                TreeRef child{smask, W, pt.cnt + G + g, nW, spx2, spm2, nsp2, nSP};
                recurse(child, m + 1);
                // pruneRedundancy then cluster, on the device in stream order
                const int *nW_dev = pt.cnt + G + g;
                if (!lab) {
                    arena.release(child_mark);
                    lab = arena.take<uint8_t>(n);  // survives until this dimension's join
                    HIP_OK(launch_fill_u8(lab, 0, n, st));
                    HIP_OK(launch_redundancy_seed(W, nW_dev, nW, stamp, ++tree_id, entries, d_nE, smask, sub.comp,
                                                  d_list, d_nlist, d_mn, prm.r, d_keep, lab, st));
                } else {
                    HIP_OK(launch_redundancy_seed(W, nW_dev, nW, stamp, ++tree_id, entries, d_nE, smask, sub.comp,
                                                  d_list, d_nlist, d_mn, prm.r, d_keep, lab, st));
                    arena.release(child_mark);
                }
                if (trace_on) {
                    HIP_OK(hipMemcpyAsync(pin, d_keep, sizeof(int), hipMemcpyDeviceToHost, st));
                    sync();
                    trace[at].redundant = pin_i()[0] ? 0 : 1;
                }
            }
            if (lab) join(smask, lab);
            arena.release(dim_mark);
        }
        arena.release(lvl_mark);
    }

    // join (util.cu:655-743), on the device
    void join(dmask_t smask, const uint8_t *lab) {
        Subspace &s = ensure_components(smask);
        HIP_OK(launch_join(lab, n, s.comp, s.csize, prm.min_size, prm.r, entries, d_nE, smask, s.entry, s.inres,
                           d_list, d_nlist, keepf, dropf, freshf, st));
    }

    // Result in the reference's order (ascending bitmask == vec_cmp,
    // util.cu:633-653) with canonical labels computed on the device, read back
    // in chunks of rows straight into the result's contiguous label array.
    void emit(inscy_result &out) {
        out.n = n;
        sync();
        std::vector<Entry> tab(nentries);
        if (nentries) HIP_OK(hipMemcpy(tab.data(), entries, (size_t) nentries * sizeof(Entry), hipMemcpyDeviceToHost));
        std::vector<int> rows;
        for (int e = 0; e < nentries; ++e)
            if (tab[e].nkept > 0) rows.push_back(e);
        std::sort(rows.begin(), rows.end(), [&](int a, int b) { return tab[a].mask < tab[b].mask; });
        const int m = (int) rows.size();
        for (int e : rows) {
            std::vector<int32_t> dims;
            for (int k = 0; k < d; ++k)
                if (dhas(tab[e].mask, k)) dims.push_back(k);
            out.subspaces.push_back(std::move(dims));
        }
        out.labels.reset(alloc_labels((size_t) std::max(m, 1) * n));
        // chunks of rows: labels on the device, read back into one of two
        // pinned buffers while the host copies the previous chunk out with
        // several threads.  A row's labels run to its kept clusters - 1, so
        // when every row has at most 255 (65,535) clusters they travel as
        // label + 1 in one (two) bytes and are widened on the host
        int maxk = 0;
        for (int e : rows) maxk = std::max(maxk, tab[e].nkept);
        const int width = maxk <= 255 ? 1 : (maxk <= 65535 ? 2 : 4);
        const size_t chunk_bytes = (size_t) 128 << 20;
        const int R = (int) std::max<size_t>(1, std::min<size_t>((size_t) m, chunk_bytes / ((size_t) n * 4)));
        if (m > 0) {
            batch.reset();
            const size_t items = (size_t) R * n + 1;
            int *flags = batch.take<int>(items);
            int *idx = batch.take<int>(items);
            char *lab = batch.take<char>((size_t) R * n * width);
            int *cnts = batch.take<int>(R);
            const size_t sb = scan_temp_bytes((int) items);
            void *tmp = batch.take<char>(sb);
            // the row counts follow the labels at a 256-byte boundary
            auto cnt_at = [&](int r) { return ((size_t) r * n * width + 255) & ~(size_t) 255; };
            const size_t stage_bytes = cnt_at(R) + (size_t) R * sizeof(int);
            char *stage[2] = {emit_stage[0].get(stage_bytes), emit_stage[1].get(stage_bytes)};
            hipEvent_t done[2];
            HIP_OK(event_pool().get(&done[0]));
            HIP_OK(event_pool().get(&done[1]));
            const unsigned nthr = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
            auto copy_out = [&](int c0, int r, const char *src) {
                int32_t *dst = out.labels.get() + (size_t) c0 * n;
                const size_t total = (size_t) r * n;
                const size_t per = (total + nthr - 1) / nthr;
                std::vector<std::thread> th;
                th.reserve(nthr);
                for (unsigned t = 0; t < nthr; ++t) {
                    const size_t lo = std::min(total, t * per), hi = std::min(total, lo + per);
                    if (lo >= hi) continue;
                    auto job = [=] {
                        if (width == 4) {
                            std::memcpy(dst + lo, reinterpret_cast<const int32_t *>(src) + lo,
                                        (hi - lo) * sizeof(int32_t));
                        } else if (width == 1) {
                            const uint8_t *s8 = reinterpret_cast<const uint8_t *>(src);
                            for (size_t i = lo; i < hi; ++i) dst[i] = (int32_t) s8[i] - 1;
                        } else {
                            const uint16_t *s16 = reinterpret_cast<const uint16_t *>(src);
                            for (size_t i = lo; i < hi; ++i) dst[i] = (int32_t) s16[i] - 1;
                        }
                    };
                    try {
                        th.emplace_back(job);
                    } catch (const std::system_error &) {
                        job();  // no thread available: this thread copies the slice
                    }
                }
                for (auto &t : th) t.join();
                const int *hc = reinterpret_cast<const int *>(src + cnt_at(r));
                for (int t = 0; t < r; ++t) {
                    if (width < 4 && hc[t] > (width == 1 ? 255 : 65535))
                        throw Failure(INSCY_ERR_DEVICE, "internal: a result row has more clusters than its table entry");
                    stats.clusters += hc[t];
                }
            };
            int prev_c0 = -1, prev_r = 0, prev_buf = 0;
            int buf = 0;
            for (int c0 = 0; c0 < m; c0 += R) {
                const int r = std::min(R, m - c0);
                std::vector<EmitRow> er(r);
                for (int t = 0; t < r; ++t) er[t] = EmitRow{tab[rows[c0 + t]].comp, tab[rows[c0 + t]].inres};
                EmitRow *der = upload(batch, er);
                HIP_OK(launch_emit(der, r, n, flags, idx, lab, width, cnts, tmp, sb, st));
                HIP_OK(hipMemcpyAsync(stage[buf], lab, (size_t) r * n * width, hipMemcpyDeviceToHost, st));
                HIP_OK(hipMemcpyAsync(stage[buf] + cnt_at(r), cnts, (size_t) r * sizeof(int),
                                      hipMemcpyDeviceToHost, st));
                HIP_OK(hipEventRecord(done[buf], st));
                if (prev_c0 >= 0) {
                    HIP_OK(hipEventSynchronize(done[prev_buf]));
                    copy_out(prev_c0, prev_r, stage[prev_buf]);
                }
                // the device reuses flags/idx/lab for the next chunk only after
                // this chunk's read-back, in stream order
                prev_c0 = c0;
                prev_r = r;
                prev_buf = buf;
                buf ^= 1;
            }
            HIP_OK(hipEventSynchronize(done[prev_buf]));
            copy_out(prev_c0, prev_r, stage[prev_buf]);
            event_pool().put(done[0]);
            event_pool().put(done[1]);
        }
        stats.subspaces = (int64_t) m;
        std::vector<unsigned long long> h2(16 * 64);
        HIP_OK(hipMemcpy(h2.data(), small + kStatAt, h2.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (int b = 0; b < 64; ++b) {
            stats.candidates += (int64_t) h2[16 * b];
            stats.queries += (int64_t) h2[16 * b + 1];
            stats.count_bytes += (int64_t) h2[16 * b + 2];
            stats.count_examined_bytes += (int64_t) h2[16 * b + 3];
        }
        // device time of the neighbour-count kernel, from the event pairs
        for (size_t i = 0; i < timers.size(); ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, timers[i].start, timers[i].stop) == hipSuccess) stats.count_kernel_ms += ms;
        }
        stats.count_launches = (int64_t) timers.size();
    }

    inscy_params prm;
    int n, d;
    hipStream_t st;
    int cells = 0, ecells = 0;
    float s_star = 0.f;
    int weak_thr = 0;
    std::vector<int> dense_thr;
    float *Xt = nullptr;
    uint8_t *cellC = nullptr;
    uint16_t *cellE = nullptr;
    dmask_t *nearm = nullptr;
    int mw = 1;  // words per path mask: 1 for d <= 64, else 2
    unsigned long long *small = nullptr;
    int *stamp = nullptr;
    int *qmark = nullptr;  // [n] stamp of the tree whose points are being counted
    int mark_id = 0;
    void *pin = nullptr;
    std::vector<double> score;  // per-dimension index selectivity (lower is better)
    long long grid_cap = 0;
    size_t scan_bytes = 0;
    DevData D{};
    Arena arena;
    Slabs slabs;
    std::unordered_map<dmask_t, Subspace, DMaskHash> subs;
    // device result table
    Entry *entries = nullptr;
    int *scratch = nullptr;
    int ecap = 0, nentries = 0;
    int *d_nE = nullptr, *d_nlist = nullptr, *d_keep = nullptr, *d_list = nullptr, *d_mn = nullptr;
    uint8_t *keepf = nullptr, *dropf = nullptr, *freshf = nullptr;
    int tree_id = 0;

    // ---------------------------------------------------------- level-synchronous driver state
    struct LTree {
        dmask_t smask = 0;
        int first = 0, m = -1, lo = 0, hi = 0, pre = 0, nW = 0;
        bool pass = false;
        const int *W = nullptr;      // device: its points (weak-dense points of its group)
        int *spx = nullptr;          // device: its s-connection paths (while its level is open)
        uint64_t *spm = nullptr;     // their masks, mw words each
        int nsp_max = 0, nsp = 0;
        int slot = -1;               // its path counter in the level's counter array
        std::vector<int> kids;       // children in (dimension, cell group) order
    };
    // one child subspace S | {m} of the frontier trees act[t0..t1) (all of mask S)
    struct ChildSub {
        dmask_t s;
        int m, t0, t1;
    };
    std::vector<LTree> lt;
    Pool store;     // weak-dense point lists of surviving trees, keep log (whole run)
    Pool lvl[2];    // frontier paths and counters of the current / next level
    Pool batch;     // one batch of child subspaces
    // grid index sorts kept for later batches, by index dimensions (whole run)
    struct KeptSort {
        int *order, *pos, *off;
    };
    std::unordered_map<uint64_t, KeptSort> kept;
    Pool keptp;
    size_t kept_used = 0, kept_budget = 0;
    // sorted coordinate columns kept the same way, by index dimensions and
    // dimension: (column, number of the batch that gathers it)
    std::unordered_map<uint64_t, std::pair<float *, int>> kept_cols;
    Pool colp;
    size_t cols_used = 0;
    int batch_no = 0;
    int cur_lvl = 0;
    HostBuf up, down;
    HostBuf emit_stage[2];
    size_t up_top = 0;
    bool up_defer = false;                          // defer_uploads()
    char *pend_dst = nullptr;                       // the deferred copy: device place,
    size_t pend_src = 0, pend_bytes = 0;            // staging offset and bytes
    size_t batch_budget = 0;
    size_t max_batch = 0;  // child subspaces per batch
    bool phi_mode = false;          // rectangular = false
    float phi_weak = 0.f;
    std::vector<float> phi_dense;   // [k]
    int *rank_perm = nullptr;       // [n] points in the neighbourhood tree's leaf order
    std::vector<int> level, lvl_of; // SCY-tree level -> dimension, dimension -> level
    int *d_level = nullptr;         // both, on the device ([level | lvl_of])
    bool batch_trace = std::getenv("INSCY_TRACE_BATCHES") != nullptr;  // one stderr line per batch
    long long pair_stats[6] = {0, 0, 0, 0, 0, 0};
    std::vector<int> lv_act, lv_next, lv_gidoff;
    std::vector<std::vector<std::vector<Group>>> lv_groups;
    int *lv_gid = nullptr;
    int *lv_next_counts = nullptr;
    int lv_nslots = 0;

    void put(void *dst, const void *src, size_t bytes);
    template <class T>
    T *upload(Pool &pool, const std::vector<T> &v) {
        T *dst = pool.take<T>(v.size());
        put(dst, v.data(), v.size() * sizeof(T));
        return dst;
    }
    void expand_level(const std::vector<int> &frontier, const int *counts);
    void process_batch(const std::vector<ChildSub> &cs, size_t i0, size_t i1);
    void replay_staged();
    std::vector<int> sup_flat;  // superset rows of every subspace (host copy of d_sup)
    static unsigned blocks_for_host(long long count) { return (unsigned) ((count + kBlockThreads - 1) / kBlockThreads); }
    // replay state: superset rows per subspace, decisions per tree
    std::vector<dmask_t> emask;                  // mask of every result-table row
    std::unordered_map<dmask_t, std::pair<int, int>, DMaskHash> sup_at;  // subspace -> (offset, count) in d_sup
    int *d_sup = nullptr;
    int *keepv = nullptr;
    int rep_join = 0;  // number of the last join
    void set_bounds(dmask_t s, BatchSub &x);
    void trace_tree(int f, const std::vector<int> &keep);
    std::vector<KernelTimer> timers;
    // pairs of the level: first pair of each active tree, pruner decisions
    std::vector<int> lv_pair0;
    std::vector<char> lv_skip;
    bool pair_on(int a, int m) const {
        return lv_skip.empty() || !lv_skip[lv_pair0[a] + m - lt[lv_act[a]].first];
    }
    void score_pairs(int na);
    // collected pairs (whole run): features, tree, dimension, children
    long long col_base = 0;
    std::vector<float> col_feat;
    std::vector<int> col_tree, col_m, col_child0, col_G;
    void finish_collect(inscy_result &out);
};

}  // namespace inscy
