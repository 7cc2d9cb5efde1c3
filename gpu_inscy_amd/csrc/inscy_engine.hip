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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/inscy.h"
#include "inscy_internal.h"

struct inscy_result {
    int64_t n = 0;
    std::vector<std::vector<int32_t>> subspaces;
    std::vector<std::vector<int32_t>> labels;
    inscy_stats stats{};
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
static float gamma_half(int m) {
    if (m == 2) return 1.0f;
    if (m == 1) return (float) std::sqrt(3.14);
    return (float) ((m / 2.0 - 1.0) * (double) gamma_half(m - 2));
}
static float ball(int k) {
    float v = (float) std::pow(3.14, k / 2.0);
    return v / gamma_half(k + 2);
}
static float expden(int k, float eps, int n) {
    float nc = (float) n * ball(k);
    float v = (float) ((double) nc * std::pow((double) eps, (double) k));
    return (float) ((double) v / std::pow(1.0, (double) k));
}
// Smallest integer count c with (float)c >= t (counts are compared as floats
// in the reference, ScyTreeNode.cu:616 and ClusteringCpu.cu:176-183).
static int count_threshold(float t) {
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
        if (base) (void) hipFree(base);
    }
    void init(size_t bytes) {
        cap = bytes;
        HIP_OK(hipMalloc(&base, bytes));
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
        for (void *p : chunks) (void) hipFree(p);
    }
    template <class T>
    T *take(size_t count) {
        size_t bytes = ((count * sizeof(T)) + 255) & ~(size_t) 255;
        if (bytes > left) {
            size_t sz = std::max(chunk, bytes);
            void *p = nullptr;
            HIP_OK(hipMalloc(&p, sz));
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

struct TreeRef {
    uint64_t smask;
    const int *P;      // device
    const int *nP;     // device count
    int nP_max;        // host bound
    const int *spx;
    const uint64_t *spm;
    const int *nsp;
    int nsp_max;
};

struct Subspace {
    int *cnt = nullptr;       // [n] neighbour counts, -1 = not computed yet
    int *comp = nullptr;      // [n] dense component root or -1
    int *csize = nullptr;     // [n] component sizes indexed by root
    uint8_t *inres = nullptr; // [n] component kept in the result
    bool comps = false;
};

class Engine {
public:
    Engine(const inscy_params &p, int n, int d, hipStream_t st) : prm(p), n(n), d(d), st(st) {}

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

        size_t nd = (size_t) n * d;
        HIP_OK(hipMalloc(&Xt, nd * sizeof(float)));
        HIP_OK(hipMalloc(&cellC, nd));
        HIP_OK(hipMalloc(&cellE, nd * sizeof(uint16_t)));
        HIP_OK(hipMalloc(&nearm, (size_t) n * sizeof(uint64_t)));
        HIP_OK(hipMalloc(&sid1, nd * sizeof(int)));
        HIP_OK(hipMalloc(&off1, (size_t) d * (ecells + 1) * sizeof(int)));
        HIP_OK(hipMalloc(&small, 64 * sizeof(unsigned long long)));
        HIP_OK(hipMalloc(&stamp, (size_t) n * sizeof(int)));
        HIP_OK(hipHostMalloc(&pin, pin_bytes()));
        // arena: per recursion depth the child lists (points, paths, survivors)
        size_t per_level = (size_t) n * (4 * sizeof(int)) + (size_t) n * 2 * (sizeof(int) + sizeof(uint64_t)) * 2;
        arena.init(per_level * (size_t) (d + 2) + table_bytes_max() + (size_t) n * 16 + (64u << 20));

        int *hist = arena.take<int>((size_t) d * (ecells + 1));
        HIP_OK(launch_fill_i32(reinterpret_cast<int *>(small), 0, 128, st));
        HIP_OK(launch_discretise(d_X, n, d, cells, prm.neighborhood_size, Xt, cellC, cellE, nearm,
                                 reinterpret_cast<int *>(small), st));
        HIP_OK(launch_build_index(cellE, n, d, ecells, hist, off1, sid1, st));
        HIP_OK(launch_fill_i32(stamp, -1, n, st));
        HIP_OK(hipMemcpyAsync(pin, small, sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        if (reinterpret_cast<int *>(pin)[0] != 0)
            throw Failure(INSCY_ERR_ARGUMENT, "X must be finite with every value in [0, 1] (normalize first)");
        arena.release(0);
        D = DevData{Xt, cellC, cellE, sid1, off1, n, d, cells, ecells};
    }

    ~Engine() {
        for (void *p : {(void *) Xt, (void *) cellC, (void *) cellE, (void *) nearm, (void *) sid1, (void *) off1,
                        (void *) small, (void *) stamp, (void *) ptrtab})
            if (p) (void) hipFree(p);
        if (pin) (void) hipHostFree(pin);
    }

    // ---------------------------------------------------------- run
    void run(inscy_result &out) {
        // root tree: all points, every near-edge chain
        int *P = arena.take<int>(n);
        int *nP = arena.take<int>(1);
        int *spx = arena.take<int>(n);
        uint64_t *spm = arena.take<uint64_t>(n);
        int *nsp = arena.take<int>(1);
        HIP_OK(launch_iota(P, n, st));
        HIP_OK(launch_fill_i32(nP, n, 1, st));
        HIP_OK(launch_fill_i32(nsp, 0, 1, st));
        HIP_OK(launch_root_paths(nearm, n, spx, spm, nsp, st));
        TreeRef root{0ull, P, nP, n, spx, spm, nsp, n};
        recurse(root, 0);
        emit(out);
    }

    inscy_stats stats{};

private:
    // ---------------------------------------------------------- helpers
    size_t pin_bytes() const { return std::max<size_t>((size_t) n * 8 + (1u << 16), 1u << 20); }
    size_t table_slots(int nW, int bits) const {
        size_t want = (size_t) std::max(1, nW) * (size_t) std::max(1, bits) * 2;
        size_t s = 1024;
        while (s < want) s <<= 1;
        return s;
    }
    size_t table_bytes_max() const { return table_slots(n, d) * sizeof(uint64_t); }

    void sync() { HIP_OK(hipStreamSynchronize(st)); }

    SubspaceArg arg_of(uint64_t smask) const {
        SubspaceArg S{};
        S.k = 0;
        for (int k = 0; k < d; ++k)
            if ((smask >> k) & 1ull) S.dims[S.k++] = k;
        return S;
    }

    Subspace &sub_of(uint64_t smask) {
        Subspace &s = subs[smask];
        if (!s.cnt) {
            s.cnt = slabs.take<int>(n);
            HIP_OK(launch_fill_i32(s.cnt, -1, n, st));
        }
        return s;
    }

    void count(uint64_t smask, Subspace &s, const int *queries, const int *nq, int nq_max) {
        SubspaceArg S = arg_of(smask);
        HIP_OK(launch_counts(D, S, queries, nq, nq_max, s_star, s.cnt, small + 8, st));
    }

    Subspace &components(uint64_t smask) {
        Subspace &s = sub_of(smask);
        if (s.comps) return s;
        size_t mk = arena.mark();
        int *q = arena.take<int>(n);
        int *nq = arena.take<int>(1);
        HIP_OK(launch_fill_i32(nq, 0, 1, st));
        HIP_OK(launch_all_queries(s.cnt, n, q, nq, st));
        count(smask, s, q, nq, n);
        s.comp = slabs.take<int>(n);
        s.csize = slabs.take<int>(n);
        int *parent = arena.take<int>(n);
        SubspaceArg S = arg_of(smask);
        HIP_OK(launch_components(D, S, s_star, s.cnt, dense_thr[S.k], parent, s.comp, s.csize, st));
        arena.release(mk);
        s.comps = true;
        stats.components++;
        return s;
    }

    // Result entries H with S ⊆ H (subspace_of, util.cu:619-631), as device
    // pointer tables; returns their count.
    int superspaces(uint64_t smask, std::vector<uint64_t> &masks) {
        masks.clear();
        for (auto &kv : result)
            if ((kv.first & smask) == smask) masks.push_back(kv.first);
        int nH = (int) masks.size();
        if (!nH) return 0;
        if (nH > ptrcap) {
            if (ptrtab) HIP_OK(hipFree(ptrtab));
            ptrcap = std::max(nH, 2 * ptrcap);
            HIP_OK(hipMalloc(&ptrtab, (size_t) 3 * ptrcap * sizeof(void *)));
        }
        void **h = reinterpret_cast<void **>(pin);
        for (int i = 0; i < nH; ++i) {
            Subspace &s = subs[masks[i]];
            h[i] = s.comp;
            h[nH + i] = s.csize;
            h[2 * nH + i] = s.inres;
        }
        HIP_OK(hipMemcpyAsync(ptrtab, h, (size_t) 3 * nH * sizeof(void *), hipMemcpyHostToDevice, st));
        return nH;
    }

    // ---------------------------------------------------------- recursion
    void recurse(const TreeRef &T, int first) {
        stats.calls++;
        for (int m = first; m < d; ++m) {
            const uint64_t smask = T.smask | (1ull << m);
            const uint8_t *cellC_m = cellC + (size_t) m * n;
            size_t dim_mark = arena.mark();

            // restrict + mergeWithNeighbors: s-connected cells of m
            int *flags = arena.take<int>(cells);
            HIP_OK(launch_fill_i32(flags, 0, cells, st));
            HIP_OK(launch_flags(T.spx, T.spm, T.nsp, T.nsp_max, cellC_m, m, flags, st));
            HIP_OK(hipMemcpyAsync(pin, flags, cells * sizeof(int), hipMemcpyDeviceToHost, st));
            sync();
            std::vector<int> flag(reinterpret_cast<int *>(pin), reinterpret_cast<int *>(pin) + cells);

            uint8_t *lab = arena.take<uint8_t>(n);  // components labelled in this dimension loop
            bool any_seed = false;
            HIP_OK(launch_fill_u8(lab, 0, n, st));

            for (int c = 0; c < cells; ++c) {
                int lo = c;
                while (flag[c] && c < cells - 1) ++c;
                int hi = c;
                size_t grp_mark = arena.mark();
                int *cnt2 = arena.take<int>(4);
                int *Pr = arena.take<int>(std::max(T.nP_max, 1));
                int *spx = arena.take<int>(std::max(T.nsp_max, 1));
                uint64_t *spm = arena.take<uint64_t>(std::max(T.nsp_max, 1));
                int *W = arena.take<int>(std::max(T.nP_max, 1));
                HIP_OK(launch_fill_i32(cnt2, 0, 4, st));
                HIP_OK(launch_filter(T.P, T.nP, T.nP_max, T.spx, T.spm, T.nsp, T.nsp_max, cellC_m, m, lo, hi, Pr,
                                     spx, spm, cnt2, st));
                Subspace &sub = sub_of(smask);
                count(smask, sub, Pr, cnt2, T.nP_max);
                const int tid = ++tree_id;
                HIP_OK(launch_weak(Pr, cnt2, T.nP_max, sub.cnt, weak_thr, W, cnt2 + 2, stamp, tid, st));
                HIP_OK(hipMemcpyAsync(pin, cnt2, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
                sync();
                const int nPr = reinterpret_cast<int *>(pin)[0];
                const int nSPr = reinterpret_cast<int *>(pin)[1];
                const int nW = reinterpret_cast<int *>(pin)[2];
                stats.trees++;
                const bool pass = nW >= prm.min_size;
                if (trace_on) trace.push_back({(long long) smask, lo, hi, nPr, nW, pass ? 1 : 0, -1});
                if (!pass) {
                    arena.release(grp_mark);
                    continue;
                }
                // s-connection paths surviving the prune
                int *spx2 = arena.take<int>(std::max(nSPr, 1));
                uint64_t *spm2 = arena.take<uint64_t>(std::max(nSPr, 1));
                int *nsp2 = cnt2 + 3;
                if (nSPr > 0) {
                    size_t tmark = arena.mark();
                    int bits = 0;
                    for (int k = m + 1; k < d; ++k) ++bits;
                    size_t slots = table_slots(nW, bits);
                    uint64_t *table = arena.take<uint64_t>(slots);
                    HIP_OK(launch_path_prune(D, smask, W, cnt2 + 2, nW, spx, spm, cnt2 + 1, nSPr, stamp, tid,
                                             small + 16, table, (uint64_t) (slots - 1), spx2, spm2, nsp2, st));
                    arena.release(tmark);
                }
                TreeRef child{smask, W, cnt2 + 2, nW, spx2, spm2, nsp2, nSPr};
                size_t at = trace.size() - 1;
                recurse(child, m + 1);
                bool keep = prune_redundancy(smask, W, cnt2 + 2, nW);
                if (trace_on) trace[at].redundant = keep ? 0 : 1;
                if (keep) {
                    Subspace &sc = components(smask);
                    HIP_OK(launch_seed(W, cnt2 + 2, nW, sc.comp, lab, st));
                    any_seed = true;
                }
                arena.release(grp_mark);
            }
            if (any_seed) join(smask, lab);
            arena.release(dim_mark);
        }
    }

    // pruneRedundancy (ScyTreeNode.cu:669-734)
    bool prune_redundancy(uint64_t smask, const int *W, const int *nW_dev, int nW) {
        std::vector<uint64_t> masks;
        int nH = superspaces(smask, masks);
        if (!nH) return true;
        size_t mk = arena.mark();
        int *mn = arena.take<int>(nH);
        HIP_OK(launch_fill_i32(mn, INT_MAX, nH, st));
        HIP_OK(launch_redundancy(W, nW_dev, nW, (int const *const *) ptrtab, (int const *const *) (ptrtab + nH),
                                 (uint8_t const *const *) (ptrtab + 2 * nH), nH, mn, st));
        HIP_OK(hipMemcpyAsync(pin, mn, nH * sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        arena.release(mk);
        int max_min = 0;
        for (int h = 0; h < nH; ++h) {
            int v = reinterpret_cast<int *>(pin)[h];
            if (v != INT_MAX && v > max_min) max_min = v;
        }
        if (max_min == 0) return true;
        float lhs = (float) nW * prm.r;
        return (double) lhs > (double) max_min;
    }

    // join (util.cu:655-743)
    void join(uint64_t smask, const uint8_t *lab) {
        Subspace &s = components(smask);
        size_t mk = arena.mark();
        int *list = arena.take<int>((size_t) 2 * n);
        int *nl = arena.take<int>(1);
        HIP_OK(launch_fill_i32(nl, 0, 1, st));
        HIP_OK(launch_collect(lab, n, s.csize, list, nl, st));
        HIP_OK(hipMemcpyAsync(pin, nl, sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        int cnt = reinterpret_cast<int *>(pin)[0];
        if (cnt == 0) {
            arena.release(mk);
            return;
        }
        HIP_OK(hipMemcpyAsync(pin, list, (size_t) 2 * cnt * sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        std::vector<int> keep;
        for (int i = 0; i < cnt; ++i) {
            int id = reinterpret_cast<int *>(pin)[2 * i], sz = reinterpret_cast<int *>(pin)[2 * i + 1];
            if (sz >= prm.min_size) keep.push_back(id);
        }
        if (keep.empty()) {
            arena.release(mk);
            return;
        }
        std::vector<uint64_t> masks;
        int nH = superspaces(smask, masks);
        if (nH) {
            uint8_t *kf = arena.take<uint8_t>(n);
            uint8_t *drop = arena.take<uint8_t>(n);
            int *ids = arena.take<int>(keep.size());
            int *hp = reinterpret_cast<int *>(pin) + (size_t) 3 * nH * 2;  // after the pointer table
            std::copy(keep.begin(), keep.end(), hp);
            HIP_OK(hipMemcpyAsync(ids, hp, keep.size() * sizeof(int), hipMemcpyHostToDevice, st));
            HIP_OK(launch_fill_u8(kf, 0, n, st));
            HIP_OK(launch_fill_u8(drop, 0, n, st));
            HIP_OK(launch_set_flags(kf, ids, (int) keep.size(), st));
            HIP_OK(launch_join_check(s.comp, kf, s.csize, n, prm.r, (int const *const *) ptrtab,
                                     (int const *const *) (ptrtab + nH), (uint8_t const *const *) (ptrtab + 2 * nH),
                                     nH, drop, st));
            sync();
            uint8_t *hd = reinterpret_cast<uint8_t *>(pin);
            HIP_OK(hipMemcpyAsync(hd, drop, n, hipMemcpyDeviceToHost, st));
            sync();
            std::vector<int> kept;
            for (int id : keep)
                if (!hd[id]) kept.push_back(id);
            keep.swap(kept);
        }
        if (!keep.empty()) {
            std::set<int> &entry = result[smask];
            if (!s.inres) {
                s.inres = slabs.take<uint8_t>(n);
                HIP_OK(launch_fill_u8(s.inres, 0, n, st));
            }
            int *ids = arena.take<int>(keep.size());
            int *hp = reinterpret_cast<int *>(pin);
            std::copy(keep.begin(), keep.end(), hp);
            HIP_OK(hipMemcpyAsync(ids, hp, keep.size() * sizeof(int), hipMemcpyHostToDevice, st));
            HIP_OK(launch_set_flags(s.inres, ids, (int) keep.size(), st));
            sync();
            entry.insert(keep.begin(), keep.end());
        }
        arena.release(mk);
    }

    void emit(inscy_result &out) {
        out.n = n;
        int *lab = arena.take<int>(n);
        for (auto &kv : result) {
            Subspace &s = subs[kv.first];
            HIP_OK(launch_labels(s.comp, s.inres, n, lab, st));
            HIP_OK(hipMemcpyAsync(pin, lab, (size_t) n * sizeof(int), hipMemcpyDeviceToHost, st));
            sync();
            // canonical labels: clusters numbered in order of their smallest point id,
            // which is the component root
            std::vector<int> roots(kv.second.begin(), kv.second.end());
            std::unordered_map<int, int> idx;
            for (size_t i = 0; i < roots.size(); ++i) idx[roots[i]] = (int) i;
            std::vector<int32_t> labels(n);
            const int *h = reinterpret_cast<int *>(pin);
            for (int i = 0; i < n; ++i) labels[i] = h[i] < 0 ? -1 : idx[h[i]];
            std::vector<int32_t> dims;
            for (int k = 0; k < d; ++k)
                if ((kv.first >> k) & 1ull) dims.push_back(k);
            out.subspaces.push_back(std::move(dims));
            out.labels.push_back(std::move(labels));
            stats.clusters += (int64_t) roots.size();
        }
        stats.subspaces = (int64_t) out.subspaces.size();
        unsigned long long h2[2];
        HIP_OK(hipMemcpyAsync(pin, small + 8, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        sync();
        memcpy(h2, pin, sizeof(h2));
        stats.candidates = (int64_t) h2[0];
        stats.queries = (int64_t) h2[1];
    }

public:
    struct TraceRow {
        long long mask;
        int lo, hi, pre, post, pass, redundant;
    };
    bool trace_on = false;
    std::vector<TraceRow> trace;

private:
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
    uint64_t *nearm = nullptr;
    int *sid1 = nullptr, *off1 = nullptr;
    unsigned long long *small = nullptr;
    int *stamp = nullptr;
    void *pin = nullptr;
    void **ptrtab = nullptr;
    int ptrcap = 0;
    DevData D{};
    Arena arena;
    Slabs slabs;
    std::unordered_map<uint64_t, Subspace> subs;
    std::map<uint64_t, std::set<int>> result;  // ascending bitmask == the reference's vec_cmp order
    int tree_id = 0;
};

static thread_local std::string g_error;
static thread_local std::vector<Engine::TraceRow> g_trace;
static bool g_trace_on = false;

static int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

static int check_params(int64_t n, int32_t d, const inscy_params *p, int32_t mode) {
    if (!p) return fail(INSCY_ERR_ARGUMENT, "params is NULL");
    if (n < 1 || n > (int64_t) INT_MAX / 2) return fail(INSCY_ERR_ARGUMENT, "n out of range");
    if (d < 1 || d > kMaxDims) return fail(INSCY_ERR_ARGUMENT, "d must be in [1, 64]");
    if (!(p->neighborhood_size > 0.0f) || !std::isfinite(p->neighborhood_size))
        return fail(INSCY_ERR_ARGUMENT, "neighborhood_size must be > 0");
    if (p->number_of_cells < 1 || p->number_of_cells > 255)
        return fail(INSCY_ERR_ARGUMENT, "number_of_cells must be in [1, 255]");
    if (p->entropy_order != 0) return fail(INSCY_ERR_UNSUPPORTED, "entropy_order != 0 is not implemented");
    if (p->rectangular == 0) return fail(INSCY_ERR_UNSUPPORTED, "rectangular=False (phi density) is not implemented");
    if (mode != INSCY_MODE_EXACT && mode != INSCY_MODE_BATCHED) return fail(INSCY_ERR_ARGUMENT, "unknown mode");
    return INSCY_OK;
}

static int run_impl(const float *d_X, int64_t n, int32_t d, const inscy_params *p, int32_t mode, hipStream_t user,
                    inscy_result **out) {
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = user;
    bool own = false;
    try {
        if (!st) {
            HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            own = true;
        }
        auto res = std::make_unique<inscy_result>();
        {
            Engine eng(*p, (int) n, d, st);
            eng.trace_on = g_trace_on;
            eng.setup(d_X);
            eng.run(*res);
            res->stats = eng.stats;
            if (g_trace_on) g_trace = eng.trace;
        }
        if (own) (void) hipStreamDestroy(st);
        res->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = res.release();
        g_error.clear();
        return INSCY_OK;
    } catch (const Failure &f) {
        if (own) (void) hipStreamDestroy(st);
        return fail(f.code, f.what());
    } catch (const std::bad_alloc &) {
        if (own) (void) hipStreamDestroy(st);
        return fail(INSCY_ERR_MEMORY, "host allocation failed");
    }
}

}  // namespace inscy

using namespace inscy;

extern "C" {

int inscy_run_device(const float *d_X, int64_t n, int32_t d, const inscy_params *params, int32_t mode, void *stream,
                     inscy_result **out) {
    if (!out) return fail(INSCY_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    int rc = check_params(n, d, params, mode);
    if (rc) return rc;
    if (!d_X) return fail(INSCY_ERR_ARGUMENT, "X is NULL");
    return run_impl(d_X, n, d, params, mode, (hipStream_t) stream, out);
}

int inscy_run(const float *X, int64_t n, int32_t d, const inscy_params *params, int32_t mode, inscy_result **out) {
    if (!out) return fail(INSCY_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    int rc = check_params(n, d, params, mode);
    if (rc) return rc;
    if (!X) return fail(INSCY_ERR_ARGUMENT, "X is NULL");
    float *dX = nullptr;
    size_t bytes = (size_t) n * d * sizeof(float);
    hipError_t e = hipMalloc(&dX, bytes);
    if (e != hipSuccess) return fail(INSCY_ERR_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemcpy(dX, X, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void) hipFree(dX);
        return fail(INSCY_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    rc = run_impl(dX, n, d, params, mode, nullptr, out);
    (void) hipFree(dX);
    return rc;
}

int32_t inscy_result_num_subspaces(const inscy_result *res) { return res ? (int32_t) res->subspaces.size() : 0; }

int32_t inscy_result_subspace(const inscy_result *res, int32_t i, int32_t *dims_out) {
    if (!res || i < 0 || i >= (int32_t) res->subspaces.size()) return -1;
    const auto &s = res->subspaces[i];
    if (dims_out) std::copy(s.begin(), s.end(), dims_out);
    return (int32_t) s.size();
}

const int32_t *inscy_result_labels(const inscy_result *res, int32_t i) {
    if (!res || i < 0 || i >= (int32_t) res->labels.size()) return nullptr;
    return res->labels[i].data();
}

int64_t inscy_result_num_points(const inscy_result *res) { return res ? res->n : 0; }

int inscy_result_stats(const inscy_result *res, inscy_stats *out) {
    if (!res || !out) return fail(INSCY_ERR_ARGUMENT, "NULL argument");
    *out = res->stats;
    return INSCY_OK;
}

void inscy_result_free(inscy_result *res) { delete res; }

const char *inscy_last_error(void) { return g_error.c_str(); }

// Test hook (not in the public header): record one trace row per evaluated
// restricted tree — (mask, lo, hi, points, survivors, pass, redundant).
int inscy_debug_trace(int on) {
    g_trace_on = on != 0;
    return INSCY_OK;
}

int64_t inscy_debug_trace_rows(int64_t *out, int64_t cap) {
    int64_t m = (int64_t) g_trace.size();
    if (out)
        for (int64_t i = 0; i < std::min(m, cap); ++i) {
            const auto &t = g_trace[i];
            int64_t *o = out + 7 * i;
            o[0] = t.mask;
            o[1] = t.lo;
            o[2] = t.hi;
            o[3] = t.pre;
            o[4] = t.post;
            o[5] = t.pass;
            o[6] = t.redundant;
        }
    return m;
}

int inscy_discretise(const float *d_X, int64_t n, int32_t d, int32_t cells, float eps, uint8_t *d_cell_c,
                     uint16_t *d_cell_e, uint64_t *d_near_mask) {
    if (!d_X || !d_cell_c || !d_cell_e || !d_near_mask || n < 1 || d < 1 || d > kMaxDims || cells < 1 ||
        cells > 255 || !(eps > 0.0f))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    try {
        int *bad = nullptr;
        HIP_OK(hipMalloc(&bad, sizeof(int)));
        HIP_OK(hipMemset(bad, 0, sizeof(int)));
        HIP_OK(launch_discretise(d_X, (int) n, d, cells, eps, nullptr, d_cell_c, d_cell_e, d_near_mask, bad, 0));
        int h = 0;
        HIP_OK(hipMemcpy(&h, bad, sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipFree(bad));
        if (h) return fail(INSCY_ERR_ARGUMENT, "X must be finite with every value in [0, 1]");
        return INSCY_OK;
    } catch (const Failure &f) {
        return fail(f.code, f.what());
    }
}

// Standalone context for the kernel-level entry points.
namespace {
struct Standalone {
    int n, d, ecells;
    float *Xt = nullptr;
    uint8_t *cellC = nullptr;
    uint16_t *cellE = nullptr;
    uint64_t *nearm = nullptr;
    int *sid1 = nullptr, *off1 = nullptr, *hist = nullptr, *bad = nullptr;
    DevData D{};
    Standalone(const float *d_X, int n_, int d_, float eps, int cells) : n(n_), d(d_) {
        ecells = (int) std::ceil(1.0 / (double) eps);
        size_t nd = (size_t) n * d;
        HIP_OK(hipMalloc(&Xt, nd * sizeof(float)));
        HIP_OK(hipMalloc(&cellC, nd));
        HIP_OK(hipMalloc(&cellE, nd * 2));
        HIP_OK(hipMalloc(&nearm, (size_t) n * 8));
        HIP_OK(hipMalloc(&sid1, nd * 4));
        HIP_OK(hipMalloc(&off1, (size_t) d * (ecells + 1) * 4));
        HIP_OK(hipMalloc(&hist, (size_t) d * (ecells + 1) * 4));
        HIP_OK(hipMalloc(&bad, 4));
        HIP_OK(hipMemset(bad, 0, 4));
        HIP_OK(launch_discretise(d_X, n, d, cells, eps, Xt, cellC, cellE, nearm, bad, 0));
        HIP_OK(launch_build_index(cellE, n, d, ecells, hist, off1, sid1, 0));
        int h = 0;
        HIP_OK(hipMemcpy(&h, bad, 4, hipMemcpyDeviceToHost));
        if (h) throw Failure(INSCY_ERR_ARGUMENT, "X must be finite with every value in [0, 1]");
        D = DevData{Xt, cellC, cellE, sid1, off1, n, d, cells, ecells};
    }
    ~Standalone() {
        for (void *p : {(void *) Xt, (void *) cellC, (void *) cellE, (void *) nearm, (void *) sid1, (void *) off1,
                        (void *) hist, (void *) bad})
            if (p) (void) hipFree(p);
    }
};

SubspaceArg make_arg(const int32_t *dims, int32_t k, int d) {
    SubspaceArg S{};
    S.k = k;
    for (int i = 0; i < k; ++i) {
        if (dims[i] < 0 || dims[i] >= d || (i && dims[i] <= dims[i - 1]))
            throw Failure(INSCY_ERR_ARGUMENT, "dims must be ascending ids in [0, d)");
        S.dims[i] = dims[i];
    }
    return S;
}
}  // namespace

int inscy_neighbour_counts(const float *d_X, int64_t n, int32_t d, float eps, const int32_t *dims, int32_t k,
                           const int32_t *d_queries, int64_t nq, int32_t *d_out) {
    if (!d_X || !dims || k < 1 || k > d || !d_queries || !d_out || n < 1 || d < 1 || d > kMaxDims || !(eps > 0.0f))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    try {
        Standalone ctx(d_X, (int) n, d, eps, 1);
        SubspaceArg S = make_arg(dims, k, d);
        int *cnt = nullptr, *nqd = nullptr;
        HIP_OK(hipMalloc(&cnt, (size_t) n * 4));
        HIP_OK(hipMalloc(&nqd, 4));
        int nqh = (int) nq;
        HIP_OK(hipMemcpy(nqd, &nqh, 4, hipMemcpyHostToDevice));
        HIP_OK(launch_fill_i32(cnt, -1, n, 0));
        HIP_OK(launch_counts(ctx.D, S, d_queries, nqd, (int) nq, sqrt_threshold(eps), cnt, nullptr, 0));
        // gather per query (queries may repeat)
        std::vector<int> hq(nq), hc(n);
        HIP_OK(hipMemcpy(hq.data(), d_queries, nq * 4, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(hc.data(), cnt, (size_t) n * 4, hipMemcpyDeviceToHost));
        std::vector<int> ho(nq);
        for (int64_t i = 0; i < nq; ++i) ho[i] = hc[hq[i]];
        HIP_OK(hipMemcpy(d_out, ho.data(), nq * 4, hipMemcpyHostToDevice));
        HIP_OK(hipFree(cnt));
        HIP_OK(hipFree(nqd));
        return INSCY_OK;
    } catch (const Failure &f) {
        return fail(f.code, f.what());
    }
}

int inscy_dense_components(const float *d_X, int64_t n, int32_t d, const inscy_params *p, const int32_t *dims,
                           int32_t k, int32_t *d_comp) {
    if (!d_X || !p || !dims || k < 1 || k > d || !d_comp || n < 1 || d < 1 || d > kMaxDims ||
        !(p->neighborhood_size > 0.0f))
        return fail(INSCY_ERR_ARGUMENT, "invalid argument");
    if (!p->rectangular) return fail(INSCY_ERR_UNSUPPORTED, "rectangular=False is not implemented");
    try {
        Standalone ctx(d_X, (int) n, d, p->neighborhood_size, std::max(1, (int) p->number_of_cells));
        SubspaceArg S = make_arg(dims, k, d);
        int *cnt = nullptr, *q = nullptr, *nqd = nullptr, *par = nullptr, *csize = nullptr;
        HIP_OK(hipMalloc(&cnt, (size_t) n * 4));
        HIP_OK(hipMalloc(&q, (size_t) n * 4));
        HIP_OK(hipMalloc(&nqd, 4));
        HIP_OK(hipMalloc(&par, (size_t) n * 4));
        HIP_OK(hipMalloc(&csize, (size_t) n * 4));
        int nn = (int) n;
        HIP_OK(hipMemcpy(nqd, &nn, 4, hipMemcpyHostToDevice));
        HIP_OK(launch_iota(q, nn, 0));
        HIP_OK(launch_fill_i32(cnt, -1, n, 0));
        float eps = p->neighborhood_size;
        HIP_OK(launch_counts(ctx.D, S, q, nqd, nn, sqrt_threshold(eps), cnt, nullptr, 0));
        int thr = count_threshold(std::max(p->F * expden(k, eps, nn), (float) p->num_obj));
        HIP_OK(launch_components(ctx.D, S, sqrt_threshold(eps), cnt, thr, par, d_comp, csize, 0));
        HIP_OK(hipDeviceSynchronize());
        for (void *ptr : {(void *) cnt, (void *) q, (void *) nqd, (void *) par, (void *) csize}) HIP_OK(hipFree(ptr));
        return INSCY_OK;
    } catch (const Failure &f) {
        return fail(f.code, f.what());
    }
}

}  // extern "C"
