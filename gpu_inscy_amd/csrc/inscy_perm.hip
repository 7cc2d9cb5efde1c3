// Permuted SCY-tree level order (entropy_order != 0, get_subspace_order,
// inscy_map.cpp:61-121): the depth-first driver over s-connection chains.
//
// With the levels in a permuted order, restrict_node (ScyTreeNode.cu:227-320)
// keeps every node whose dimension id exceeds the restricted one, so a
// restricted tree holds empty regular nodes and truncated s-connection chains,
// and mergeNodes (ScyTreeNode.cu:803-833) adds counts and points only into
// nodes whose count is positive.  A tree here is its points P (the weak-dense
// points of its last prune) and its chains: (x, b, end) is the chain point x
// starts at the level of dimension b, running through x's cells down to level
// `end` (restricted levels skipped).  Per restriction on m and cell group G:
//  - merge flags: chains with b == m, by x's cell in m;
//  - points: those with their cell in G; when m's level is the deepest
//    unrestricted one and the level above it has a dimension id above m (or
//    none is left), only prefixes holding a point of G's first cell keep
//    theirs (the tree of that cell has an empty node there, which mergeNodes
//    does not add to);
//  - chains (b != m): a chain below m's level is reached only through x's
//    cell and keeps its nodes down to the deepest one with dimension id above
//    m; a chain above passes when its cell is in G (keeping the part below
//    down to the deepest such node, else ending just above m's level), and
//    otherwise keeps its nodes down to the deepest one above m's level with a
//    dimension id above m;
//  - prune (pruneRecursionAndRemove2 + propergate_count2): a chain stays when
//    its regular parent keeps a weak-dense point, unless its same-cell
//    regular sibling existed before the prune (a prefix of the tree's points,
//    an ancestor of one of its chains, or a node above m's level kept because
//    a level between it and m's has a dimension id above m) and lost its
//    weak-dense points.
// Prefix membership tests use open-addressing tables of (32 hash bits |
// representative point), every hit confirmed by comparing cells.
#include "inscy_engine.h"

#include "inscy_device.h"

namespace inscy {

namespace {

__device__ __forceinline__ int chain_b(uint64_t w) { return (int) (w & 0xFFull); }
__device__ __forceinline__ int chain_end(uint64_t w) { return (int) ((w >> 8) & 0xFFull); }
__device__ __forceinline__ uint64_t chain_word(int b, int end) { return (uint64_t) b | ((uint64_t) end << 8); }

constexpr uint64_t kSeed = 0x243F6A8885A308D3ull;
__device__ __forceinline__ uint64_t level_salt(int k) { return 0x9E3779B97F4A7C15ull * (uint64_t) (k + 1); }

// prefix of z through level index `upto` (unrestricted levels of smask)
__device__ __forceinline__ uint64_t prefix_hash(const DevData &D, dmask_t smask, int z, int upto) {
    uint64_t h = kSeed;
    for (int t = 0; t <= upto; ++t) {
        const int k = D.level[t];
        if (dhas(smask, k)) continue;
        h = mix64(h + D.cellC[(size_t) k * D.n + z] + 1);
    }
    return mix64(h ^ level_salt(D.level[upto]));
}

__device__ void table_insert(const DevData &D, dmask_t smask, uint64_t *table, uint64_t tmask, uint64_t e, int z,
                             int t) {
    const uint64_t entry = (e & 0xFFFFFFFF00000000ull) | (uint32_t) z;
    uint64_t slot = e & tmask;
    while (true) {
        uint64_t cur = table[slot];
        if (cur == kEmpty) {
            cur = atomicCAS((unsigned long long *) &table[slot], (unsigned long long) kEmpty,
                            (unsigned long long) entry);
            if (cur == kEmpty) return;
        }
        if ((cur >> 32) == (entry >> 32) && same_prefix(D, smask, z, (int) (uint32_t) cur, t)) return;
        slot = (slot + 1) & tmask;
    }
}

__device__ bool table_has(const DevData &D, dmask_t smask, const uint64_t *table, uint64_t tmask, uint64_t e, int z,
                          int t) {
    uint64_t slot = e & tmask;
    while (true) {
        const uint64_t cur = table[slot];  // complete: the inserts ran in an earlier launch
        if (cur == kEmpty) return false;
        if ((cur >> 32) == (e >> 32) && same_prefix(D, smask, z, (int) (uint32_t) cur, t)) return true;
        slot = (slot + 1) & tmask;
    }
}

__global__ void k_perm_root(const dmask_t *__restrict__ nearm, int n, int d, int *__restrict__ cx,
                            uint64_t *__restrict__ cm, int *__restrict__ nout) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    dmask_t m = nearm[p];
    if (!m) return;
    int at = atomicAdd(nout, dcount(m));
    while (m) {
        const int b = dfirst(m);
        m &= m - 1;
        cx[at] = p;
        cm[at] = chain_word(b, d - 1);
        ++at;
    }
}

__global__ void k_perm_flags(const int *__restrict__ cx, const uint64_t *__restrict__ cm, const int *__restrict__ nch,
                             const uint8_t *__restrict__ cellC, int n, int first, int cells, int *__restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nch) return;
    const int b = chain_b(cm[i]);
    if (b < first) return;
    flags[(size_t) (b - first) * cells + cellC[(size_t) b * n + cx[i]]] = 1;
}

// the chains of group g after the restriction of T (smask) on m
__global__ void k_perm_route(DevData D, dmask_t smask, int m, const int *__restrict__ gid, int g,
                             const int *__restrict__ cx, const uint64_t *__restrict__ cm, const int *__restrict__ nch,
                             int *__restrict__ ox, uint64_t *__restrict__ om, int *__restrict__ nout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nch) return;
    const uint64_t w = cm[i];
    const int b = chain_b(w), end = chain_end(w);
    if (b == m) return;
    const int x = cx[i];
    const bool cell_in = gid[D.cellC[(size_t) m * D.n + x]] == g;
    const int Lb = D.lvl_of[b], Lm = D.lvl_of[m];
    auto deepest = [&](int from, int to) {  // deepest unrestricted level in [from, to], not m's, dimension > m
        int best = -1;
        for (int t = from; t <= to; ++t) {
            const int k = D.level[t];
            if (t != Lm && !dhas(smask, k) && k > m) best = t;
        }
        return best;
    };
    int e;
    if (Lb > Lm) {
        if (!cell_in) return;
        e = deepest(Lb, end);
    } else if (end >= Lm && cell_in) {
        e = deepest(Lm + 1, end);
        if (e < 0) {
            e = Lm - 1;  // the chain node above the restricted level
            while (e > Lb && dhas(smask, D.level[e])) --e;
        }
    } else {
        e = deepest(Lb, min(end, Lm - 1));
    }
    if (e < 0) return;
    const int at = atomicAdd(nout, 1);
    ox[at] = x;
    om[at] = chain_word(b, e);
}

// prefixes of the points at every unrestricted level whose dimension is in `need`
__global__ void k_perm_insert_pts(DevData D, dmask_t smask, const int *__restrict__ pts, const int *__restrict__ npts,
                                  dmask_t need, uint64_t *__restrict__ table, uint64_t tmask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *npts || !need) return;
    const int z = pts[i];
    uint64_t h = kSeed;
    for (int t = 0; t < D.d; ++t) {
        const int k = D.level[t];
        if (dhas(smask, k)) continue;
        h = mix64(h + D.cellC[(size_t) k * D.n + z] + 1);
        if (dhas(need, k)) table_insert(D, smask, table, tmask, mix64(h ^ level_salt(k)), z, t);
    }
}

// regular ancestors of the chains' first nodes: x's prefixes at the levels above b's
__global__ void k_perm_insert_anchors(DevData D, dmask_t smask, const int *__restrict__ cx,
                                      const uint64_t *__restrict__ cm, const int *__restrict__ nch,
                                      uint64_t *__restrict__ table, uint64_t tmask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nch) return;
    const int z = cx[i], Lb = D.lvl_of[chain_b(cm[i])];
    uint64_t h = kSeed;
    for (int t = 0; t < Lb; ++t) {
        const int k = D.level[t];
        if (dhas(smask, k)) continue;
        h = mix64(h + D.cellC[(size_t) k * D.n + z] + 1);
        table_insert(D, smask, table, tmask, mix64(h ^ level_salt(k)), z, t);
    }
}

struct PermTables {
    const uint64_t *W, *P, *T, *A;
    uint64_t wmask, pmask, tmask, amask;
    dmask_t cond;  // dimensions whose nodes above m's level a larger dimension id below keeps
};

__global__ void k_perm_prune(DevData D, dmask_t smask, PermTables Tb, const int *__restrict__ cx,
                             const uint64_t *__restrict__ cm, const int *__restrict__ nch, int *__restrict__ ox,
                             uint64_t *__restrict__ om, int *__restrict__ nout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nch) return;
    const int z = cx[i];
    const uint64_t w = cm[i];
    const int b = chain_b(w), Lb = D.lvl_of[b];
    // the regular parent: x's prefix through the last unrestricted level above b's
    int tp = Lb - 1;
    while (tp >= 0 && dhas(smask, D.level[tp])) --tp;
    if (tp >= 0 && !table_has(D, smask, Tb.W, Tb.wmask, prefix_hash(D, smask, z, tp), z, tp)) return;
    const uint64_t e = prefix_hash(D, smask, z, Lb);
    if (!table_has(D, smask, Tb.W, Tb.wmask, e, z, Lb)) {
        const bool existed = table_has(D, smask, Tb.P, Tb.pmask, e, z, Lb) ||
                             table_has(D, smask, Tb.A, Tb.amask, e, z, Lb) ||
                             (dhas(Tb.cond, b) && table_has(D, smask, Tb.T, Tb.tmask, e, z, Lb));
        if (existed) return;  // the sibling lost its points: propergate_count2 erases the chain
    }
    const int at = atomicAdd(nout, 1);
    ox[at] = z;
    om[at] = w;
}

// the points of G's first cell mark their prefixes through level index tkey
// (tkey < 0: the root, flag[0])
__global__ void k_perm_lo_mark(DevData D, dmask_t smask, const int *__restrict__ pts, const int *__restrict__ npts,
                               int m, int lo, int tkey, uint64_t *__restrict__ table, uint64_t tmask,
                               int *__restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *npts) return;
    const int z = pts[i];
    if (D.cellC[(size_t) m * D.n + z] != lo) return;
    if (tkey < 0)
        flag[0] = 1;
    else
        table_insert(D, smask, table, tmask, prefix_hash(D, smask, z, tkey), z, tkey);
}

__global__ void k_perm_lo_filter(DevData D, dmask_t smask, const int *__restrict__ in, const int *__restrict__ nin,
                                 int tkey, const uint64_t *__restrict__ table, uint64_t tmask,
                                 const int *__restrict__ flag, int *__restrict__ out, int *__restrict__ nout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nin) return;
    const int z = in[i];
    const bool keep = tkey < 0 ? flag[0] != 0 : table_has(D, smask, table, tmask, prefix_hash(D, smask, z, tkey), z, tkey);
    if (keep) out[atomicAdd(nout, 1)] = z;
}

#define LAUNCH(kernel, grid, block, lds, st, ...) hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__)

size_t slots_for(long long items) {
    size_t want = (size_t) std::max<long long>(items, 1) * 2, s = 256;
    while (s < want) s <<= 1;
    return s;
}

}  // namespace

void Engine::run_perm(inscy_result &out) {
    int *P = arena.take<int>(n);
    int *nP = arena.take<int>(1);
    const size_t cap = (size_t) n * d;
    int *cx = arena.take<int>(std::max<size_t>(cap, 1));
    uint64_t *cm = arena.take<uint64_t>(std::max<size_t>(cap, 1));
    int *nch = arena.take<int>(1);
    HIP_OK(launch_iota(P, n, st));
    HIP_OK(launch_fill_i32(nP, n, 1, st));
    HIP_OK(launch_fill_i32(nch, 0, 1, st));
    LAUNCH(k_perm_root, dim3(blocks_for(n)), dim3(kBlock), 0, st, nearm, n, d, cx, cm, nch);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(pin, nch, sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    const int nch_h = pin_i()[0];
    TreeRef root{(dmask_t) 0, P, nP, n, cx, nullptr, nch, nch_h, cm};
    auto t0 = std::chrono::steady_clock::now();
    recurse_perm(root, 0);
    sync();
    auto t1 = std::chrono::steady_clock::now();
    emit(out);
    stats.expand_seconds = std::chrono::duration<double>(t1 - t0).count();
    stats.emit_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
}

void Engine::recurse_perm(const TreeRef &T, int first) {
    stats.calls++;
    if (first >= d) return;
    const int nd = d - first;
    const size_t lvl_mark = arena.mark();
    int *flags = arena.take<int>((size_t) nd * cells);
    HIP_OK(launch_fill_i32(flags, 0, (size_t) nd * cells, st));
    if (T.nsp_max > 0)
        LAUNCH(k_perm_flags, dim3(blocks_for(T.nsp_max)), dim3(kBlock), 0, st, T.spx, T.chain, T.nsp, cellC, n, first,
               cells, flags);
    HIP_OK(hipGetLastError());
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

    for (int m = first; m < d; ++m) {
        const dmask_t smask = T.smask | dbit(m);
        const auto &gs = groups[m - first];
        const int G = (int) gs.size();
        const size_t dim_mark = arena.mark();
        int *cnt = arena.take<int>((size_t) 6 * G);
        int *PP = arena.take<int>(std::max(T.nP_max, 1));
        int *PW = arena.take<int>(std::max(T.nP_max, 1));
        const Subspace &sub = subs[smask];
        HIP_OK(launch_partition(T.P, T.nP, T.nP_max, nullptr, nullptr, mw, nullptr, 0, cellC + (size_t) m * n,
                                d_gid + (size_t) (m - first) * cells, G, sub.cnt, weak_thr, phi_mode ? 1 : 0, m, cnt, PP,
                                PW, nullptr, nullptr, st));
        HIP_OK(hipMemcpyAsync(pin, cnt, (size_t) 2 * G * sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        std::vector<int> c2(pin_i(), pin_i() + 2 * G);
        // m's level the deepest unrestricted one, and the level above it:
        // the point filter of merged groups
        const int Lm = lvl_of[m];
        int deeper = 0, tabove = -1;
        for (int t = 0; t < d; ++t) {
            const int k = level[t];
            if (dhas(T.smask, k) || k == m) continue;
            if (t > Lm) deeper = 1;
            else tabove = t;
        }
        const bool filt = !deeper && (tabove < 0 || level[tabove] > m);
        // dimensions of nodes above m's level kept empty by a larger dimension id below
        dmask_t cond = 0;
        for (int t = 0; t < Lm; ++t) {
            const int k = level[t];
            if (dhas(T.smask, k)) continue;
            for (int u = t; u < Lm; ++u)
                if (!dhas(T.smask, level[u]) && level[u] > m) cond |= dbit(k);
        }
        // the dimension's cluster marks, below every group's buffers (a
        // group's buffers are released while its kernels may still be queued:
        // later launches on the stream reuse the memory only after them)
        uint8_t *lab_buf = arena.take<uint8_t>(n);
        uint8_t *lab = nullptr;
        int offP = 0, offW = 0;
        for (int g = 0; g < G; ++g) {
            int nPr = c2[g], nW = c2[G + g];
            const int *Pg = PP + offP, *Wg = PW + offW;
            offP += nPr;
            offW += nW;
            const size_t group_mark = arena.mark();
            int *nPg = arena.take<int>(2);
            int *nWg = nPg + 1;
            HIP_OK(launch_fill_i32(nPg, nPr, 1, st));
            HIP_OK(launch_fill_i32(nWg, nW, 1, st));
            if (filt && gs[g].hi > gs[g].lo && nPr > 0) {
                const size_t slots = slots_for(nPr);
                uint64_t *table = arena.take<uint64_t>(slots);
                int *flag = arena.take<int>(3);
                int *P2 = arena.take<int>(std::max(nPr, 1));
                int *W2 = arena.take<int>(std::max(nW, 1));
                HIP_OK(launch_fill_u64(table, kEmpty, slots, st));
                HIP_OK(launch_fill_i32(flag, 0, 3, st));
                LAUNCH(k_perm_lo_mark, dim3(blocks_for(nPr)), dim3(kBlock), 0, st, D, smask, Pg, nPg, m, gs[g].lo,
                       tabove, table, (uint64_t) (slots - 1), flag);
                LAUNCH(k_perm_lo_filter, dim3(blocks_for(nPr)), dim3(kBlock), 0, st, D, smask, Pg, nPg, tabove,
                       table, (uint64_t) (slots - 1), flag, P2, flag + 1);
                if (nW > 0)
                    LAUNCH(k_perm_lo_filter, dim3(blocks_for(nW)), dim3(kBlock), 0, st, D, smask, Wg, nWg, tabove,
                           table, (uint64_t) (slots - 1), flag, W2, flag + 2);
                HIP_OK(hipGetLastError());
                HIP_OK(hipMemcpyAsync(pin, flag + 1, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
                sync();
                nPr = pin_i()[0];
                nW = pin_i()[1];
                Pg = P2;
                Wg = W2;
                HIP_OK(launch_fill_i32(nPg, nPr, 1, st));
                HIP_OK(launch_fill_i32(nWg, nW, 1, st));
            }
            stats.trees++;
            const bool pass = nW >= prm.min_size;
            const size_t at = trace.size();
            if (trace_on) trace.push_back({(long long) dlo(smask), gs[g].lo, gs[g].hi, nPr, nW, pass ? 1 : 0, -1});
            if (!pass) {
                arena.release(group_mark);
                continue;
            }
            // the group's chains, then their prune
            const int cap = std::max(T.nsp_max, 1);
            int *rx = arena.take<int>(cap);
            uint64_t *rm = arena.take<uint64_t>(cap);
            int *nr = arena.take<int>(1);
            int *cx2 = arena.take<int>(cap);
            uint64_t *cm2 = arena.take<uint64_t>(cap);
            int *nc2 = arena.take<int>(1);
            HIP_OK(launch_fill_i32(nr, 0, 1, st));
            HIP_OK(launch_fill_i32(nc2, 0, 1, st));
            if (T.nsp_max > 0) {
                LAUNCH(k_perm_route, dim3(blocks_for(T.nsp_max)), dim3(kBlock), 0, st, D, T.smask, m,
                       d_gid + (size_t) (m - first) * cells, g, T.spx, T.chain, T.nsp, rx, rm, nr);
                int free_dims = 0;
                dmask_t need = 0;
                for (int k = 0; k < d; ++k)
                    if (!dhas(smask, k)) {
                        ++free_dims;
                        need |= dbit(k);
                    }
                const size_t ws = slots_for((long long) nW * free_dims), ps = slots_for((long long) nPr * free_dims),
                             ts = slots_for((long long) T.nP_max * free_dims),
                             as = slots_for((long long) T.nsp_max * free_dims);
                uint64_t *tw = arena.take<uint64_t>(ws), *tp = arena.take<uint64_t>(ps), *tt = arena.take<uint64_t>(ts),
                         *ta = arena.take<uint64_t>(as);
                HIP_OK(launch_fill_u64(tw, kEmpty, ws, st));
                HIP_OK(launch_fill_u64(tp, kEmpty, ps, st));
                HIP_OK(launch_fill_u64(tt, kEmpty, ts, st));
                HIP_OK(launch_fill_u64(ta, kEmpty, as, st));
                if (nW > 0)
                    LAUNCH(k_perm_insert_pts, dim3(blocks_for(nW)), dim3(kBlock), 0, st, D, smask, Wg, nWg, need, tw,
                           (uint64_t) (ws - 1));
                if (nPr > 0)
                    LAUNCH(k_perm_insert_pts, dim3(blocks_for(nPr)), dim3(kBlock), 0, st, D, smask, Pg, nPg, need, tp,
                           (uint64_t) (ps - 1));
                if (cond && T.nP_max > 0)
                    LAUNCH(k_perm_insert_pts, dim3(blocks_for(T.nP_max)), dim3(kBlock), 0, st, D, smask, T.P, T.nP,
                           cond & need, tt, (uint64_t) (ts - 1));
                LAUNCH(k_perm_insert_anchors, dim3(blocks_for(T.nsp_max)), dim3(kBlock), 0, st, D, smask, rx, rm, nr,
                       ta, (uint64_t) (as - 1));
                const PermTables Tb{tw, tp, tt, ta, ws - 1, ps - 1, ts - 1, as - 1, cond};
                LAUNCH(k_perm_prune, dim3(blocks_for(T.nsp_max)), dim3(kBlock), 0, st, D, smask, Tb, rx, rm, nr, cx2,
                       cm2, nc2);
                HIP_OK(hipGetLastError());
            }
            Subspace &s2 = ensure_components(smask);
            TreeRef child{smask, Wg, nWg, nW, cx2, nullptr, nc2, T.nsp_max, cm2};
            recurse_perm(child, m + 1);
            // pruneRedundancy then cluster, on the device in stream order
            if (!lab) {
                lab = lab_buf;
                HIP_OK(launch_fill_u8(lab, 0, n, st));
            }
            HIP_OK(launch_redundancy_seed(Wg, nWg, nW, stamp, ++tree_id, entries, d_nE, smask, s2.comp, d_list,
                                          d_nlist, d_mn, prm.r, d_keep, lab, st));
            if (trace_on) {
                HIP_OK(hipMemcpyAsync(pin, d_keep, sizeof(int), hipMemcpyDeviceToHost, st));
                sync();
                trace[at].redundant = pin_i()[0] ? 0 : 1;
            }
            arena.release(group_mark);
        }
        if (lab) join(smask, lab);
        arena.release(dim_mark);
    }
    arena.release(lvl_mark);
}

}  // namespace inscy
