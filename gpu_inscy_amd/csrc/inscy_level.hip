// Level-synchronous driver of the MI355X INSCY engine (INSCY_MODE_BATCHED).
//
// In the reference's depth-first recursion (INSCYCPU3Weak,
// src/algorithms/inscy/InscyNodeCpu.cu:159-235) only pruneRedundancy and join
// read the result built so far.  restrict, mergeWithNeighbors, the
// weak-density prune and the s-connection prune read nothing but the parent
// tree, and the dense components of a subspace read nothing but the data.  So
// the tree of restricted trees is grown here one level at a time:
//   - all trees of a level get their merge flags in one pass and one read-back;
//   - the level's child subspaces go through in batches: one grid-index build,
//     one neighbour-count pass over the trees' points, one partition pass of
//     every (tree, dimension) pair and one read-back of the group sizes;
//   - the surviving children of a batch get their path prune, and their new
//     subspaces their dense components, a handful of launches for all of them.
// pruneRedundancy and join are then replayed on the device in stages that keep
// the reference's depth-first order wherever it matters (steps on subspaces
// one of which contains the other), which gives the depth-first driver's
// result exactly.

#include "inscy_engine.h"

#include <functional>

namespace inscy {

namespace {
constexpr size_t kMaxBatch = 1024;  // child subspaces per batch (inscy_set_tuning, INSCY_MAX_BATCH override)
constexpr uint64_t kEmptySlot = ~0ull;  // free slot of a path-prune table
inline int nblocks(long long count) { return (int) ((count + kBlockThreads - 1) / kBlockThreads); }
}  // namespace

void Engine::put(void *dst, const void *src, size_t bytes) {
    if (!bytes) return;
    const size_t need = (bytes + 255) & ~(size_t) 255;
    if (up_top + need > up.cap) {
        const bool defer = up_defer;
        if (up_top > 0) sync();  // staged copies still read the buffer (sync issues a deferred one)
        up_defer = defer;
        up_top = 0;
        if (need > up.cap) up.get(std::max(need, 2 * up.cap));
    }
    std::memcpy(up.p + up_top, src, bytes);
    if (up_defer) {
        // pool takes and staging places are both rounded to 256 bytes, so a
        // contiguous pair extends the pending copy (padding bytes included:
        // they lie inside the take)
        if (pend_bytes && static_cast<char *>(dst) == pend_dst + pend_bytes && up_top == pend_src + pend_bytes) {
            pend_bytes += need;
        } else {
            if (pend_bytes) HIP_OK(hipMemcpyAsync(pend_dst, up.p + pend_src, pend_bytes, hipMemcpyHostToDevice, st));
            pend_dst = static_cast<char *>(dst);
            pend_src = up_top;
            pend_bytes = need;
        }
    } else {
        HIP_OK(hipMemcpyAsync(dst, up.p + up_top, bytes, hipMemcpyHostToDevice, st));
    }
    up_top += need;
}

void Engine::run_levels(inscy_result &out) {
    lt.clear();
    emask.clear();
    col_base = 0;
    col_feat.clear();
    col_tree.clear();
    col_m.clear();
    col_child0.clear();
    col_G.clear();
    store.chunk = (size_t) 256 << 20;
    batch.chunk = (size_t) 256 << 20;
    up.get((size_t) 4 << 20);
    max_batch = kMaxBatch;
    if (const char *e = std::getenv("INSCY_MAX_BATCH")) max_batch = (size_t) std::max(1, std::atoi(e));
    if (max_batch_override > 0) max_batch = (size_t) max_batch_override;
    batch_budget = (size_t) 2 << 30;
    if (const size_t free_b = device_available_bytes() / (size_t) mem_share)
        batch_budget = std::min<size_t>(std::max<size_t>(free_b / 8, (size_t) 256 << 20), (size_t) 8 << 30);
    kept.clear();
    keptp.reset();
    kept_used = 0;
    kept_budget = keep_budget_override >= 0 ? (size_t) keep_budget_override : batch_budget;
    kept_cols.clear();
    colp.reset();
    cols_used = 0;
    batch_no = 0;

    // the root tree: every point, the root s-connection paths
    Pool &L0 = lvl[cur_lvl];
    lvl[cur_lvl ^ 1].reset();
    int *P = store.take<int>(n);
    HIP_OK(launch_iota(P, n, st));
    int *cnt0 = L0.take<int>(1);
    LTree root;
    root.first = 0;
    root.pass = true;
    root.W = P;
    root.nW = n;
    root.pre = n;
    root.spx = L0.take<int>(n);
    root.spm = L0.take<uint64_t>((size_t) n * mw);
    root.nsp_max = n;
    root.slot = 0;
    HIP_OK(launch_root_paths(nearm, n, root.spx, root.spm, mw, cnt0, st));
    lt.push_back(std::move(root));
    stats.calls = 1;

    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> frontier{0};
    const int *counts = cnt0;
    while (!frontier.empty()) {
        expand_level(frontier, counts);
        frontier.swap(lv_next);
        counts = lv_next_counts;
        lvl[cur_lvl].reset();
        cur_lvl ^= 1;
    }

    // children in the reference's visiting order: dimension, then cell group
    for (auto &t : lt)
        std::stable_sort(t.kids.begin(), t.kids.end(), [&](int a, int b) {
            return lt[a].m != lt[b].m ? lt[a].m < lt[b].m : lt[a].lo < lt[b].lo;
        });
    sync();
    auto t1 = std::chrono::steady_clock::now();
    // superset rows of every subspace with a row: AND of per-dimension bitsets
    {
        const int ne = (int) emask.size();
        const int words = (ne + 63) / 64;
        std::vector<std::vector<uint64_t>> has(d, std::vector<uint64_t>(words, 0));
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < d; ++k)
                if (dhas(emask[e], k)) has[k][e >> 6] |= 1ull << (e & 63);
        std::vector<int> flat;
        std::vector<uint64_t> acc(words);
        sup_at.clear();
        for (int e = 0; e < ne; ++e) {
            const dmask_t sm = emask[e];
            std::fill(acc.begin(), acc.end(), ~0ull);
            for (int k = 0; k < d; ++k)
                if (dhas(sm, k))
                    for (int w = 0; w < words; ++w) acc[w] &= has[k][w];
            const int at = (int) flat.size();
            for (int w = 0; w < words; ++w)
                for (uint64_t b = acc[w]; b; b &= b - 1) flat.push_back(w * 64 + __builtin_ctzll(b));
            sup_at[sm] = {at, (int) flat.size() - at};
        }
        d_sup = upload(store, flat);
        sup_flat.swap(flat);
    }
    keepv = store.take<int>(lt.size());
    rep_join = 0;
    replay_staged();
    if (collect) finish_collect(out);
    sync();
    auto t2 = std::chrono::steady_clock::now();
    if (trace_on) {
        std::vector<int> keep(lt.size());
        HIP_OK(hipMemcpyAsync(keep.data(), keepv, keep.size() * sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        trace.clear();
        trace_tree(0, keep);
    }
    auto t3 = std::chrono::steady_clock::now();
    if (batch_trace)
        std::fprintf(stderr, "[inscy pairs] %lld (tree, dimension) pairs, %lld with a surviving child; "
                             "points counted %lld, in pairs with a surviving child %lld; pairs whose largest "
                             "cell group holds fewer than min_size points %lld (points %lld)\n",
                     pair_stats[0], pair_stats[1], pair_stats[2], pair_stats[3], pair_stats[4], pair_stats[5]);
    emit(out);
    stats.expand_seconds = std::chrono::duration<double>(t1 - t0).count();
    stats.replay_seconds = std::chrono::duration<double>(t2 - t1).count();
    stats.emit_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t3).count();
}

// One level: merge flags and exact path counts of the frontier, then the
// child subspaces in batches.  Fills lv_next (surviving children that can
// still restrict) and lv_next_counts (their path counters).
void Engine::expand_level(const std::vector<int> &frontier, const int *counts) {
    Pool &cur = lvl[cur_lvl];
    Pool &nxt = lvl[cur_lvl ^ 1];
    lv_next.clear();
    lv_nslots = 0;
    lv_next_counts = nullptr;
    lv_act.clear();
    for (int t : frontier)
        if (lt[t].first < d) lv_act.push_back(t);
    if (lv_act.empty()) return;
    const int na = (int) lv_act.size();

    // merge flags of every dimension >= first of every tree
    std::vector<FlagTree> ft(na);
    std::vector<int> fblk(na + 1, 0);
    lv_gidoff.assign(na, 0);
    int ftot = 0;
    for (int a = 0; a < na; ++a) {
        const LTree &t = lt[lv_act[a]];
        ft[a] = FlagTree{t.spx, t.spm, counts + t.slot, t.first, ftot, mw};
        lv_gidoff[a] = ftot;
        ftot += (d - t.first) * cells;
        fblk[a + 1] = fblk[a] + nblocks(t.nsp_max);
    }
    int *flags = cur.take<int>(ftot);
    HIP_OK(launch_fill_i32(flags, 0, ftot, st));
    defer_uploads();
    FlagTree *dft = upload(cur, ft);
    int *dfb = upload(cur, fblk);
    flush_uploads();
    HIP_OK(launch_b_flags(dft, dfb, na, fblk[na], cellC, n, cells, flags, st));
    const int nf = (int) frontier.size();
    char *h = down.get((size_t) (ftot + nf) * sizeof(int));
    HIP_OK(hipMemcpyAsync(h, flags, (size_t) ftot * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(h + (size_t) ftot * sizeof(int), counts, (size_t) nf * sizeof(int), hipMemcpyDeviceToHost,
                          st));
    sync();
    const int *hf = reinterpret_cast<const int *>(h);
    const int *hc = hf + ftot;
    int max_slot = -1;
    for (int t : frontier) max_slot = std::max(max_slot, lt[t].slot);
    if (max_slot >= nf) throw Failure(INSCY_ERR_DEVICE, "internal: path counter slot out of range");
    for (int t : frontier) lt[t].nsp = std::min(hc[lt[t].slot], lt[t].nsp_max);

    // cell groups (restrict + mergeWithNeighbors, ScyTreeNode.cu:837-851)
    lv_groups.assign(na, {});
    std::vector<int> gid(ftot);
    long long max_children = 0;
    for (int a = 0; a < na; ++a) {
        const LTree &t = lt[lv_act[a]];
        lv_groups[a].resize(d - t.first);
        for (int m = t.first; m < d; ++m) {
            const int *f = hf + lv_gidoff[a] + (m - t.first) * cells;
            int *gm = gid.data() + lv_gidoff[a] + (m - t.first) * cells;
            auto &gs = lv_groups[a][m - t.first];
            for (int c = 0; c < cells; ++c) {
                int lo = c;
                while (f[c] && c < cells - 1) ++c;
                for (int x = lo; x <= c; ++x) gm[x] = (int) gs.size();
                gs.push_back({lo, c});
            }
            max_children += (long long) gs.size();
        }
    }
    lv_gid = upload(cur, gid);
    lv_skip.clear();
    if (pruner || collect) score_pairs(na);
    lv_next_counts = nxt.take<int>(std::max<long long>(max_children, 1));
    HIP_OK(launch_fill_i32(lv_next_counts, 0, std::max<long long>(max_children, 1), st));

    // child subspaces: per run of frontier trees with one mask, S | {m} for m >= first
    std::vector<ChildSub> cs;
    std::set<dmask_t> seen;
    for (int a = 0; a < na;) {
        int b = a;
        while (b < na && lt[lv_act[b]].smask == lt[lv_act[a]].smask) ++b;
        const LTree &t = lt[lv_act[a]];
        // the trees of one mask are created together, so each child subspace
        // is reached from exactly one run
        if (!seen.insert(t.smask).second) throw Failure(INSCY_ERR_DEVICE, "internal: frontier masks not grouped");
        for (int m = t.first; m < d; ++m) {
            bool any = false;  // some pair of this child subspace is not skipped
            for (int x = a; x < b && !any; ++x) any = pair_on(x, m);
            if (any) cs.push_back({t.smask | dbit(m), m, a, b});
        }
        a = b;
    }

    // batches bounded by memory, count and 32-bit positions
    size_t i = 0;
    while (i < cs.size()) {
        size_t j = i;
        size_t bytes = 0;
        long long hsum = 1;
        while (j < cs.size() && j - i < max_batch) {
            const ChildSub &c = cs[j];
            SubspaceArg S = arg_of(c.s);
            GridArg G = grid_of(S, -1, true);
            const long long ncell = grid_cells(G);
            // process_batch's takes per subspace: counts, keys, rank, order,
            // query flags and list, two sample edges, packed cells (own and
            // shared), dense flags and parents (50 bytes a point), k coordinate
            // columns and cells (6k), the phi pass's queue (8), grid offsets
            size_t est = (size_t) n * (50 + 6 * (size_t) S.k + (phi_mode ? 8 : 0)) + (size_t) ncell * 8;
            for (int a = c.t0; a < c.t1; ++a) {
                if (!pair_on(a, c.m)) continue;
                const LTree &t = lt[lv_act[a]];
                est += (size_t) t.nW * (8 + 32 * (size_t) (d - 1 - c.m)) + (size_t) t.nsp * 12 + 6 * 256 * 4;
            }
            const bool fits = bytes + est <= batch_budget && (long long) (j - i + 1) * n < INT_MAX &&
                              hsum + ncell < INT_MAX;
            if (j > i && !fits) break;
            bytes += est;
            hsum += ncell;
            ++j;
        }
        process_batch(cs, i, j);
        i = j;
    }
}

void Engine::process_batch(const std::vector<ChildSub> &cs, size_t i0, size_t i1) {
    Pool &nxt = lvl[cur_lvl ^ 1];
    batch.reset();
    const int B = (int) (i1 - i0);
    const int QPB = kQueriesPerBlock;

    // ---- grid indexes of the batch's subspaces
    // The children S + {m} of one frontier run whose S has at least three
    // matched dimensions share one grid index over S's most selective ones
    // (m, the last dimension, is then checked per candidate) and the sorted
    // coordinates of S's dimensions; each adds its own column for m.
    std::vector<BatchSub> hs(B);
    std::vector<int> reps;
    long long coord = 0, hoff = 0, qoff = 0;
    int kmax = 1, kmin = 64;
    bool all_packed = true;
    int builder = -1;  // the index builder the next child of the same run may share
    std::vector<int> shared_of(B, -1);  // the builder whose index a child shares
    int *cntB = batch.take<int>((size_t) B * n);
    for (int b = 0; b < B; ++b) {
        const ChildSub &c = cs[i0 + b];
        BatchSub &x = hs[b];
        std::memset(&x, 0, sizeof(x));
        x.S = arg_of(c.s);
        const int kS = x.S.k - 1;
        // shared where it costs the neighbour counts little: S's index
        // dimensions are at most 25% less selective than the best index of
        // S + {m} (expected candidates per query scale with the product of the
        // index dimensions' selectivity scores)
        bool shareable = kS >= 3 && __builtin_popcountll(match_of(x.S) & ((1ull << kS) - 1)) >= 3;
        GridArg gbest{};
        if (shareable) {
            gbest = grid_of(x.S, x.S.k - 1, true);
            const GridArg gopt = grid_of(x.S, -1, true);
            double cs_ = 1.0, co_ = 1.0;
            for (int a = 0; a < gbest.j; ++a) cs_ *= std::max(score[x.S.dims[gbest.tI[a]]], 1.0);
            for (int a = 0; a < gopt.j; ++a) co_ *= std::max(score[x.S.dims[gopt.tI[a]]], 1.0);
            shareable = gbest.j == gopt.j && cs_ <= 1.25 * co_;
        }
        long long ncell = 1;
        if (shareable && builder >= 0 && cs[i0 + builder].t0 == c.t0 && cs[i0 + builder].t1 == c.t1 &&
            hs[builder].G.j == gbest.j &&
            std::equal(gbest.tI, gbest.tI + gbest.j, hs[builder].G.tI)) {
            const BatchSub &r = hs[builder];
            shared_of[b] = builder;
            x.G = r.G;
            x.obase = r.obase;
            x.hoff = r.hoff;
            x.coord = r.coord;
            x.pos_m = x.S.k - 1;
            x.mcoord = coord;
            x.rep = 0;
            coord += n;
            ncell = grid_cells(x.G);
        } else {
            x.G = shareable ? grid_of(x.S, x.S.k - 1, true) : grid_of(x.S, -1, true);
            ncell = grid_cells(x.G);
            x.obase = (long long) reps.size() * n;
            x.hoff = hoff;
            hoff += ncell;
            x.coord = coord;
            x.pos_m = shareable ? x.S.k - 1 : x.S.k;
            x.mcoord = shareable ? coord + (long long) (x.S.k - 1) * n : 0;
            x.rep = 1;
            coord += (long long) x.S.k * n;
            reps.push_back(b);
            if (shareable) builder = b;  // a later child of the same run may share it
        }
        long long qmax = 0;
        for (int a = c.t0; a < c.t1; ++a)
            if (pair_on(a, c.m)) qmax += lt[lv_act[a]].nW;
        x.base = (long long) b * n;
        x.qoff = qoff;
        x.cnt = cntB + (size_t) b * n;
        x.ncell = (int) ncell;
        x.cap = cap_of(x.S.k);
        x.thr_dense = dense_thr[x.S.k];
        x.qmax = (int) qmax;
        set_bounds(c.s, x);
        x.dbit = phi_mode ? 2 : 0;
        x.matchmask = match_of(x.S);
        x.cellE = D.cellE;
        x.phi_w = phi_weak;
        x.phi_d = phi_dense[x.S.k];
        x.nchk4 = 0;
        x.restmask = 0;
        for (int t = 0; t < x.S.k; ++t) {
            if (!((x.G.checkmask >> t) & 1ull)) continue;
            if (x.nchk4 < 4)
                x.chk[x.nchk4++] = t;
            else
                x.restmask |= 1ull << t;
        }
        qoff += qmax;
        kmax = std::max(kmax, x.S.k);
        kmin = std::min(kmin, x.S.k);
        all_packed = all_packed && x.restmask == 0;
    }
    const int kfast = (!generic_walk && kmin == kmax && kmax <= 7 && all_packed) ? kmax : 0;

    // index sorts kept from earlier batches of the run: the same index
    // dimensions (in index order) sort the points the same way.  The indexes
    // still to sort take the first places (obase, hoff), so the histogram
    // scan covers only them; kept sorts are copied in behind.  A sort kept in
    // this batch serves a later index of the same batch too.
    std::vector<int> miss, hit_reps;
    std::vector<KeptSort> hit_src, store_src;
    int max_cell = 0;
    auto cells_of = [&](const BatchSub &x) { return (int) grid_cells(x.G); };
    for (int r : reps) {
        const BatchSub &x = hs[r];
        const int ncell = cells_of(x);
        max_cell = std::max(max_cell, ncell);
        uint64_t key = (uint64_t) x.G.j;
        for (int a = 0; a < x.G.j; ++a) key |= (uint64_t) (x.S.dims[x.G.tI[a]] + 1) << (8 * (a + 1));
        auto it = kept.find(key);
        if (it != kept.end()) {
            hit_reps.push_back(r);
            hit_src.push_back(it->second);
            continue;
        }
        miss.push_back(r);
        const size_t need = ((size_t) 2 * n + ncell + 1) * sizeof(int) + 768;
        KeptSort ks{nullptr, nullptr, nullptr};
        if (kept_used + need <= kept_budget) {
            ks = KeptSort{keptp.take<int>(n), keptp.take<int>(n), keptp.take<int>((size_t) ncell + 1)};
            kept.emplace(key, ks);
            kept_used += need;
        }
        store_src.push_back(ks);
    }
    long long hmiss = 0;
    {
        long long slot = 0, hpos = 0;
        for (int r : miss) {
            hs[r].obase = slot++ * n;
            hs[r].hoff = hpos;
            hpos += cells_of(hs[r]);
        }
        hmiss = hpos;
        for (int r : hit_reps) {
            hs[r].obase = slot++ * n;
            hs[r].hoff = hpos;
            hpos += cells_of(hs[r]);
        }
        for (int b = 0; b < B; ++b)
            if (shared_of[b] >= 0) {
                hs[b].obase = hs[shared_of[b]].obase;
                hs[b].hoff = hs[shared_of[b]].hoff;
            }
    }
    std::vector<SortCopy> hits, stores;
    for (size_t u = 0; u < miss.size(); ++u)
        if (store_src[u].order)
            stores.push_back({store_src[u].order, store_src[u].pos, store_src[u].off, hs[miss[u]].obase,
                              hs[miss[u]].hoff, cells_of(hs[miss[u]])});
    for (size_t u = 0; u < hit_reps.size(); ++u)
        hits.push_back({hit_src[u].order, hit_src[u].pos, hit_src[u].off, hs[hit_reps[u]].obase,
                        hs[hit_reps[u]].hoff, cells_of(hs[hit_reps[u]])});
    stats.sorts_kept += (int64_t) hits.size();
    reps = miss;
    reps.insert(reps.end(), hit_reps.begin(), hit_reps.end());
    BatchIndex X{};
    X.nhist = (int) (hoff + 1);
    X.keys = batch.take<int>((size_t) B * n);
    X.rank = batch.take<int>((size_t) B * n);
    X.order = batch.take<int>((size_t) B * n + 4);  // the grouped link walk reads up to 3 ids past a run
    X.hist = batch.take<int>(X.nhist);
    X.off = batch.take<int>(X.nhist);
    X.Ys = batch.take<float>(coord + 4);  // the grouped walks read up to 3 past a column's end
    X.Es = batch.take<uint16_t>(coord);
    X.qflag = batch.take<uint8_t>(((size_t) B * n + 15) & ~(size_t) 15);  // read 16 bytes a lane
    X.qlist = batch.take<int>(std::max<long long>(qoff, 1));
    X.nsel = batch.take<int>(1);
    X.qstart = batch.take<int>(B + 1);
    X.select_bytes = select_temp_bytes(B * n);
    X.select_temp = batch.take<char>(X.select_bytes);
    X.nb = batch.take<int>((size_t) kSampleEdges * B * n);
    X.sig = batch.take<uint64_t>((size_t) B * n + 4);
    X.sigS = batch.take<uint64_t>((size_t) reps.size() * n);
    X.reps = upload(batch, reps);
    X.nrep = (int) reps.size();
    X.nfb = batch.take<int>(1);
    X.fbcap = phi_mode ? B * n : 1;  // each point of each subspace is sent at most once per pass
    X.fb = batch.take<int2>(X.fbcap);
    HIP_OK(launch_fill_i32(X.nb, -1, (size_t) kSampleEdges * B * n, st));
    X.dsorted = batch.take<uint8_t>((size_t) B * n + 8);  // the grouped link walk reads aligned words past the end
    X.par = batch.take<int>((size_t) B * n);
    X.giant = batch.take<int>(B);
    X.scan_bytes = scan_temp_bytes(X.nhist);
    X.scan_temp = batch.take<char>(X.scan_bytes);
    HIP_OK(launch_fill_i32(cntB, -1, (size_t) B * n, st));
    BatchSub *dsubs = upload(batch, hs);
    // sorted columns kept from earlier batches: the same index dimensions and
    // dimension give the same column, in the order of the index's kept sort,
    // so only an index whose sort is kept (copied in, or kept in this batch)
    // keeps or reads columns: an index sorted again orders the points of a
    // cell differently.  A column kept in this batch is read only from the
    // next one on (the batch's gathers run in one launch).
    ++batch_no;
    auto sort_key = [&](const BatchSub &x) {
        uint64_t key = (uint64_t) x.G.j;
        for (int a = 0; a < x.G.j; ++a) key |= (uint64_t) (x.S.dims[x.G.tI[a]] + 1) << (8 * (a + 1));
        return key;
    };
    auto column = [&](const BatchSub &x, int t, const float *&src, float *&dst) {
        if (!kept.count(sort_key(x))) return;
        const uint64_t ck = sort_key(x) | (uint64_t) (x.S.dims[t] + 1) << 32;
        auto it = kept_cols.find(ck);
        if (it != kept_cols.end()) {
            if (it->second.second < batch_no) src = it->second.first;
        } else if (cols_used + (size_t) n * sizeof(float) + 256 <= kept_budget) {
            dst = colp.take<float>(n);
            kept_cols.emplace(ck, std::make_pair(dst, batch_no));
            cols_used += (size_t) n * sizeof(float) + 256;
        }
    };
    std::vector<const float *> ysrc(reps.size() * kmax, nullptr), own_src(B, nullptr);
    std::vector<float *> ystore(reps.size() * kmax, nullptr), own_store(B, nullptr);
    for (size_t u = 0; u < reps.size(); ++u) {
        const BatchSub &x = hs[reps[u]];
        for (int t = 0; t < x.S.k; ++t)
            if (t != x.pos_m) column(x, t, ysrc[u * kmax + t], ystore[u * kmax + t]);
    }
    for (int b = 0; b < B; ++b)
        if (hs[b].pos_m < hs[b].S.k) column(hs[b], hs[b].pos_m, own_src[b], own_store[b]);
    // the index tables of the batch in one copy
    std::vector<char> blob;
    auto add = [&blob](const auto &v) {
        const size_t at = (blob.size() + 255) & ~(size_t) 255;
        const size_t bytes = v.size() * sizeof(v[0]);
        blob.resize(at + bytes);
        if (bytes) std::memcpy(blob.data() + at, v.data(), bytes);
        return at;
    };
    const size_t o_ysrc = add(ysrc), o_ystore = add(ystore), o_osrc = add(own_src), o_ostore = add(own_store);
    const size_t o_miss = add(miss), o_stores = add(stores), o_hits = add(hits);
    char *dblob = upload(batch, blob);
    ColumnKeep cols{reinterpret_cast<const float *const *>(dblob + o_ysrc), reinterpret_cast<float *const *>(dblob + o_ystore),
                    reinterpret_cast<const float *const *>(dblob + o_osrc),
                    reinterpret_cast<float *const *>(dblob + o_ostore), kmax};
    const int *dmiss = reinterpret_cast<const int *>(dblob + o_miss);
    const SortCopy *dstores = reinterpret_cast<const SortCopy *>(dblob + o_stores);
    const SortCopy *dhits = reinterpret_cast<const SortCopy *>(dblob + o_hits);
    HIP_OK(launch_b_index(D, dsubs, B, X, kfast == 0, dmiss, (int) miss.size(), (int) hmiss, dstores,
                          (int) stores.size(), dhits, (int) hits.size(), max_cell, cols, st));


    // ---- queries: the points of the trees each subspace refines, in sorted order
    std::vector<QuerySeg> qs;
    std::vector<int> qblk{0};
    for (int b = 0; b < B; ++b) {
        const ChildSub &c = cs[i0 + b];
        for (int a = c.t0; a < c.t1; ++a) {
            const LTree &t = lt[lv_act[a]];
            if (t.nW <= 0 || !pair_on(a, c.m)) continue;
            qs.push_back({t.W, t.nW, b});
            qblk.push_back(qblk.back() + nblocks(t.nW));
        }
    }
    defer_uploads();
    QuerySeg *dqs = upload(batch, qs);
    int *dqb = upload(batch, qblk);
    flush_uploads();
    HIP_OK(launch_b_queries(dsubs, B, n, dqs, dqb, (int) qs.size(), qblk.back(), X, st));
    std::vector<int> cblk(B + 1, 0);
    for (int b = 0; b < B; ++b) cblk[b + 1] = cblk[b] + (hs[b].qmax + QPB - 1) / QPB;
    int *dcb = upload(batch, cblk);
    const bool dbg = batch_trace;
    std::chrono::steady_clock::time_point tq0, tq1;
    if (dbg) {
        sync();
        tq0 = std::chrono::steady_clock::now();
    }
    HIP_OK(launch_b_count(dsubs, B, dcb, cblk[B], n, kmax, kfast, X, s_star, prm.neighborhood_size, phi_mode, phi_all_exact, D, rank_perm,
                          small + kStatAt, timer(), st));
    if (dbg) {
        sync();
        tq1 = std::chrono::steady_clock::now();
    }

    // ---- partition of every (tree, dimension) pair by the cell groups
    struct PairInfo {
        int a, b, G, cnt_off;
    };
    std::vector<PairInfo> pinfo;
    int pc_total = 0;
    for (int b = 0; b < B; ++b) {
        const ChildSub &c = cs[i0 + b];
        for (int a = c.t0; a < c.t1; ++a) {
            if (!pair_on(a, c.m)) continue;
            const LTree &t = lt[lv_act[a]];
            const int G = (int) lv_groups[a][c.m - t.first].size();
            pinfo.push_back({a, b, G, pc_total});
            pc_total += 6 * G;
        }
    }
    const int np = (int) pinfo.size();
    int *pcnt = batch.take<int>(std::max(pc_total, 1));
    HIP_OK(launch_fill_i32(pcnt, 0, std::max(pc_total, 1), st));
    std::vector<PartPair> pairs(np);
    std::vector<int> pblk(np + 1, 0);
    for (int j = 0; j < np; ++j) {
        const PairInfo &pi = pinfo[j];
        const ChildSub &c = cs[i0 + pi.b];
        const LTree &t = lt[lv_act[pi.a]];
        PartPair &pp = pairs[j];
        std::memset(&pp, 0, sizeof(pp));
        pp.P = t.W;
        pp.nP = t.nW;
        pp.spx = t.spx;
        pp.spm = t.spm;
        pp.nsp = t.nsp;
        pp.gid = lv_gid + lv_gidoff[pi.a] + (c.m - t.first) * cells;
        pp.weakcnt = hs[pi.b].cnt;
        pp.cnt = pcnt + pi.cnt_off;
        pp.outW = batch.take<int>(std::max(t.nW, 1));
        pp.outx = batch.take<int>(std::max(t.nsp, 1));
        pp.outm = batch.take<uint64_t>((size_t) std::max(t.nsp, 1) * mw);
        pp.mw = mw;
        pp.m = c.m;
        pp.G = pi.G;
        pp.bp = nblocks((t.nW + kPartPerThread - 1) / kPartPerThread);
        pblk[j + 1] = pblk[j] + pp.bp + nblocks((t.nsp + kPartPerThread - 1) / kPartPerThread);
    }
    defer_uploads();
    PartPair *dpairs = upload(batch, pairs);
    int *dpb = upload(batch, pblk);
    flush_uploads();
    HIP_OK(launch_b_partition(dpairs, dpb, np, pblk[np], cellC, n, weak_thr, phi_mode ? 1 : 0, prm.min_size, st));
    char *h = down.get((size_t) std::max(pc_total, 1) * sizeof(int));
    HIP_OK(hipMemcpyAsync(h, pcnt, (size_t) std::max(pc_total, 1) * sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    const std::vector<int> hcnt(reinterpret_cast<const int *>(h), reinterpret_cast<const int *>(h) + pc_total);

    // ---- children: trace rows, surviving trees, their point lists and path prunes
    std::vector<CopySeg> copies;
    std::vector<CopySeg> stamps;  // W points -> tree index, into the subspace's stamp array
    std::vector<int> stamp_b;
    std::vector<PruneChild> prune;
    std::vector<size_t> prune_slots;
    std::vector<char> passing(B, 0);
    size_t table_total = 0;
    for (int j = 0; j < np; ++j) {
        const PairInfo &pi = pinfo[j];
        const ChildSub &c = cs[i0 + pi.b];
        const int tparent = lv_act[pi.a];
        const int first = lt[tparent].first;
        const int G = pi.G;
        if (batch_trace) {
            bool any = false, big = false;
            for (int g = 0; g < G; ++g) {
                any |= hcnt[pi.cnt_off + G + g] >= prm.min_size;
                big |= hcnt[pi.cnt_off + g] >= prm.min_size;
            }
            pair_stats[0] += 1;
            pair_stats[1] += any;
            pair_stats[2] += lt[tparent].nW;
            pair_stats[3] += any ? lt[tparent].nW : 0;
            pair_stats[4] += big ? 0 : 1;
            pair_stats[5] += big ? 0 : lt[tparent].nW;
        }
        if (collect) {
            const long long gp = col_base + lv_pair0[pi.a] + (c.m - first);
            col_child0[gp] = (int) lt.size();
            col_G[gp] = G;
        }
        const int *c3 = hcnt.data() + pi.cnt_off;
        int offW = 0, offS = 0;
        for (int g = 0; g < G; ++g) {
            const int nPr = c3[g], nW = c3[G + g], nSP = c3[2 * G + g];
            LTree ch;
            ch.smask = c.s;
            ch.first = c.m + 1;
            ch.m = c.m;
            ch.lo = lv_groups[pi.a][c.m - first][g].lo;
            ch.hi = lv_groups[pi.a][c.m - first][g].hi;
            ch.pre = nPr;
            ch.nW = nW;
            ch.pass = nW >= prm.min_size;
            stats.trees++;
            if (ch.pass) {
                stats.calls++;
                passing[pi.b] = 1;
                int *Wst = store.take<int>(std::max(nW, 1));
                if (nW > 0) {
                    copies.push_back({pairs[j].outW + offW, Wst, nW, 0});
                    stamps.push_back({pairs[j].outW + offW, nullptr, nW, (int) lt.size()});
                    stamp_b.push_back(pi.b);
                }
                ch.W = Wst;
                if (ch.first < d) {
                    ch.slot = lv_nslots++;
                    ch.nsp_max = nSP;
                    ch.spx = nxt.take<int>(std::max(nSP, 1));
                    ch.spm = nxt.take<uint64_t>((size_t) std::max(nSP, 1) * mw);
                    if (nSP > 0) {
                        PruneChild pc;
                        std::memset(&pc, 0, sizeof(pc));
                        pc.smask = c.s;
                        const size_t slots = table_slots(nW, d - 1 - c.m);
                        pc.tmask = (uint64_t) (slots - 1);
                        pc.W = pairs[j].outW + offW;
                        pc.spx = pairs[j].outx + offS;
                        pc.spm = pairs[j].outm + (size_t) offS * mw;
                        pc.mw = mw;
                        pc.outx = ch.spx;
                        pc.outm = ch.spm;
                        pc.nout = lv_next_counts + ch.slot;
                        pc.nW = nW;
                        pc.nsp = nSP;
                        prune_slots.push_back(table_total);
                        table_total += slots;
                        prune.push_back(pc);
                    }
                }
            }
            const int idx = (int) lt.size();
            const bool expand = ch.pass && ch.first < d;
            lt.push_back(std::move(ch));
            lt[tparent].kids.push_back(idx);
            if (expand) lv_next.push_back(idx);
            offW += nW;
            offS += nSP;
        }
    }

    // ---- path prune of the surviving children
    if (!prune.empty()) {
        const int nc = (int) prune.size();
        uint64_t *tables = batch.take<uint64_t>(table_total);
        dmask_t *bits = batch.take<dmask_t>(nc);
        HIP_OK(launch_fill_u64(tables, kEmptySlot, table_total, st));
        HIP_OK(launch_fill_u64(reinterpret_cast<uint64_t *>(bits), 0ull, (size_t) 2 * nc, st));
        std::vector<int> bw(nc + 1, 0), bp(nc + 1, 0);
        for (int c = 0; c < nc; ++c) {
            prune[c].table = tables + prune_slots[c];
            prune[c].bits = bits + c;
            bw[c + 1] = bw[c] + nblocks(prune[c].nW);
            bp[c + 1] = bp[c] + nblocks(prune[c].nsp);
        }
        defer_uploads();
        PruneChild *dpc = upload(batch, prune);
        int *dbw = upload(batch, bw);
        int *dbp = upload(batch, bp);
        flush_uploads();
        HIP_OK(launch_b_prune(D, dpc, nc, dbw, bw[nc], dbp, bp[nc], st));
    }

    // ---- dense components of the subspaces that gained a surviving tree
    std::vector<BatchSub> hc;
    std::vector<Entry> ents;
    int ncomp = 0;
    for (int b = 0; b < B; ++b) ncomp += passing[b];
    if (ncomp > 0) {
        int *cntP = slabs.take<int>((size_t) ncomp * n);
        int *compP = slabs.take<int>((size_t) ncomp * n);
        int *csizeP = slabs.take<int>((size_t) ncomp * n);
        int *memP = slabs.take<int>((size_t) ncomp * n);
        uint8_t *inresP = slabs.take<uint8_t>((size_t) ncomp * n);
        int *stampP = slabs.take<int>((size_t) ncomp * n);
        HIP_OK(launch_fill_u8(inresP, 0, (size_t) ncomp * n, st));
        HIP_OK(launch_fill_i32(stampP, -1, (size_t) ncomp * n, st));
        std::vector<int *> stamp_of(B, nullptr);
        grow_entries(nentries + ncomp);
        int ci = 0;
        for (int b = 0; b < B; ++b) {
            if (!passing[b]) continue;
            const dmask_t s = cs[i0 + b].s;
            BatchSub x = hs[b];
            x.cnt = cntP + (size_t) ci * n;
            x.comp = compP + (size_t) ci * n;
            x.csize = csizeP + (size_t) ci * n;
            copies.push_back({hs[b].cnt, x.cnt, n, 0});
            hc.push_back(x);
            Subspace &sb = subs[s];
            sb.cnt = x.cnt;
            sb.comp = x.comp;
            sb.csize = x.csize;
            sb.inres = inresP + (size_t) ci * n;
            sb.tstamp = stampP + (size_t) ci * n;
            stamp_of[b] = sb.tstamp;
            sb.entry = nentries + ci;
            emask.push_back(s);
            ents.push_back(Entry{s, sb.comp, sb.csize, sb.inres, memP + (size_t) ci * n, 0, 0});
            ++ci;
        }
        for (size_t u = 0; u < stamps.size(); ++u) stamps[u].dst = stamp_of[stamp_b[u]];
    }
    if (!copies.empty()) {
        std::vector<int> cb(copies.size() + 1, 0);
        for (size_t c = 0; c < copies.size(); ++c) cb[c + 1] = cb[c] + nblocks(copies[c].len);
        defer_uploads();
        CopySeg *dcs = upload(batch, copies);
        int *dcb2 = upload(batch, cb);
        flush_uploads();
        HIP_OK(launch_b_copy(dcs, dcb2, (int) copies.size(), cb.back(), st));
    }
    if (!stamps.empty()) {
        std::vector<int> sbk(stamps.size() + 1, 0);
        for (size_t c = 0; c < stamps.size(); ++c) sbk[c + 1] = sbk[c] + nblocks(stamps[c].len);
        defer_uploads();
        CopySeg *dss = upload(batch, stamps);
        int *dsb = upload(batch, sbk);
        flush_uploads();
        HIP_OK(launch_b_stamp(dss, dsb, (int) stamps.size(), sbk.back(), st));
    }
    if (ncomp > 0) {
        BatchSub *dhc = upload(batch, hc);
        std::chrono::steady_clock::time_point tc0, tc1, tc2;
        if (dbg) {
            sync();
            tc0 = std::chrono::steady_clock::now();
        }
        HIP_OK(launch_b_count(dhc, ncomp, nullptr, 0, n, kmax, kfast, X, s_star, prm.neighborhood_size, phi_mode, phi_all_exact, D,
                              rank_perm, small + kStatAt, timer(), st));
        if (dbg) {
            sync();
            tc1 = std::chrono::steady_clock::now();
        }
        HIP_OK(launch_b_components(dhc, ncomp, n, kmax, kfast, X, s_star, st));
        if (dbg) {
            sync();
            tc2 = std::chrono::steady_clock::now();
            int kmin = 64, kmx = 0;
            long long qs = 0;
            for (int b = 0; b < B; ++b) {
                kmin = std::min(kmin, hs[b].S.k);
                kmx = std::max(kmx, hs[b].S.k);
                qs += hs[b].qmax;
            }
            auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
                return std::chrono::duration<double, std::milli>(b - a).count();
            };
            std::fprintf(stderr,
                         "[inscy batch] B=%d k=%d..%d cap=%d..%d queries<=%lld list_count=%.3fms "
                         "comp=%d all_count=%.3fms components=%.3fms\n",
                         B, kmin, kmx, cap_of(kmx), cap_of(kmin), qs, ms(tq0, tq1), ncomp, ms(tc0, tc1),
                         ms(tc1, tc2));
        }
        Entry *de = upload(batch, ents);
        HIP_OK(launch_b_entries(entries, nentries, de, ncomp, d_nE, st));
        nentries += ncomp;
        stats.components += ncomp;
    }
}

// The subspaces S minus one dimension whose counts are complete (their
// dense components were built) bound the counts of S from above; a bound
// decides a point only below both thresholds of S and below the cap of the
// smaller subspaces, under which their counts are exact.
void Engine::set_bounds(dmask_t s, BatchSub &x) {
    const int k = x.S.k;
    int lo = std::min(weak_thr, dense_thr[k]);
    x.nbound = 0;
    x.lo = lo;
    for (int t = 0; t < k && x.nbound < kMaxBounds; ++t) {
        const dmask_t sub = s & ~dbit(x.S.dims[t]);
        auto it = subs.find(sub);
        if (sub && it != subs.end() && it->second.entry >= 0) x.bound[x.nbound++] = it->second.cnt;
    }
    // counts are exact below the cap of |S| - 1 (phi: the stored values carry
    // an upper bound, used in the kernel)
    if (x.nbound > 0 && !phi_mode) lo = std::min(lo, cap_of(k - 1));
    x.lo = lo;
}

// pruneRedundancy and join, replayed in stages.  The reference's order visits
// per dimension every surviving child's subtree, then its pruneRedundancy,
// then the join of the kept children's clusters.  A step on subspace S reads
// the result rows of the supersets of S and writes the row of S, so two steps
// must keep their order exactly when one subspace contains the other; each
// step goes to the stage after the last earlier step it depends on, and the
// steps of one stage run together, a handful of launches per stage.
void Engine::replay_staged() {
    struct Step {
        dmask_t s;
        int kind, tree, t_lo, t_hi;
    };
    std::vector<Step> steps;
    std::vector<int> pass;
    std::function<void(int)> walk = [&](int f) {
        const std::vector<int> &kids = lt[f].kids;
        size_t i = 0;
        while (i < kids.size()) {
            const int m = lt[kids[i]].m;
            size_t j = i;
            int t_lo = INT_MAX, t_hi = -1;
            std::vector<int> pk;
            while (j < kids.size() && lt[kids[j]].m == m) {
                t_lo = std::min(t_lo, kids[j]);
                t_hi = std::max(t_hi, kids[j] + 1);
                if (lt[kids[j]].pass) pk.push_back(kids[j]);
                ++j;
            }
            const dmask_t s = lt[kids[i]].smask;
            if (pk.size() == 1) {
                walk(pk[0]);
                steps.push_back({s, 0, pk[0], pk[0], pk[0] + 1});
            } else if (!pk.empty()) {
                for (int c : pk) {
                    walk(c);
                    steps.push_back({s, 1, c, c, c + 1});
                }
                steps.push_back({s, 2, -1, t_lo, t_hi});
            }
            i = j;
        }
    };
    walk(0);
    // stages: last_on[r] = last stage of a step on row r, last_sub[r] = last
    // stage of a step on a subspace of row r
    const int ne = (int) emask.size();
    std::vector<int> last_on(ne, 0), last_sub(ne, 0), stage(steps.size());
    int nst = 0;
    for (size_t i = 0; i < steps.size(); ++i) {
        const Subspace &sb = subs.at(steps[i].s);
        const auto sl = sup_at.at(steps[i].s);
        int st0 = last_sub[sb.entry];
        for (int u = 0; u < sl.second; ++u) st0 = std::max(st0, last_on[sup_flat[sl.first + u]]);
        const int sti = st0 + 1;
        stage[i] = sti;
        nst = std::max(nst, sti);
        last_on[sb.entry] = std::max(last_on[sb.entry], sti);
        for (int u = 0; u < sl.second; ++u) {
            int &x = last_sub[sup_flat[sl.first + u]];
            x = std::max(x, sti);
        }
    }
    std::vector<std::vector<int>> by_stage(nst + 1);
    for (size_t i = 0; i < steps.size(); ++i) by_stage[stage[i]].push_back((int) i);
    if (batch_trace)
        std::fprintf(stderr, "[inscy replay] %zu steps in %d stages\n", steps.size(), nst);

    // per step: row flags (2 per superset row) and a block counter
    // (redundancy), the join's two words and its candidate list
    const size_t kScratchInts = (size_t) 64 << 20;  // per launch group
    const int ccap = prm.min_size > 0 ? std::min(n, n / prm.min_size + 1) : n;
    for (int sg = 1; sg <= nst; ++sg) {
        const std::vector<int> &ids = by_stage[sg];
        size_t a = 0;
        while (a < ids.size()) {
            // a chunk of the stage whose scratch fits the budget
            size_t b = a, ints = 0;
            while (b < ids.size()) {
                const Step &x = steps[ids[b]];
                const size_t need = 2 * (size_t) sup_at.at(x.s).second + 3 + (x.kind == 1 ? 0 : (size_t) ccap);
                if (b > a && ints + need > kScratchInts) break;
                ints += need;
                ++b;
            }
            batch.reset();
            int *scr = batch.take<int>(std::max<size_t>(ints, 1));
            HIP_OK(launch_fill_i32(scr, 0, std::max<size_t>(ints, 1), st));
            std::vector<RepOp> ops(b - a);
            std::vector<int> br{0}, bc{0}, bk{0}, bm{0};
            size_t at = 0;
            for (size_t u = a; u < b; ++u) {
                const Step &x = steps[ids[u]];
                const Subspace &sb = subs.at(x.s);
                const auto sl = sup_at.at(x.s);
                RepOp &o = ops[u - a];
                std::memset(&o, 0, sizeof(o));
                o.list = d_sup + sl.first;
                o.nlist = sl.second;
                o.tstamp = sb.tstamp;
                o.flag = sb.cnt;
                o.comp = sb.comp;
                o.csize = sb.csize;
                o.inres = sb.inres;
                o.entry = sb.entry;
                o.kind = x.kind;
                o.tree = x.tree;
                o.t_lo = x.t_lo;
                o.t_hi = x.t_hi;
                o.J = x.kind == 1 ? 0 : ++rep_join;
                o.ccap = ccap;
                o.rowflag = scr + at;
                at += 2 * (size_t) sl.second;
                o.done = reinterpret_cast<unsigned *>(scr + at);
                at += 1;
                o.words = scr + at;
                at += 2;
                if (x.kind != 1) {
                    o.clist = scr + at;
                    at += ccap;
                }
                if (x.tree >= 0) {
                    o.W = lt[x.tree].W;
                    o.nW = lt[x.tree].nW;
                }
                br.push_back(br.back() + (x.kind != 2 ? rep_red_blocks(o.nlist, o.nW) : 0));
                bc.push_back(bc.back() + (x.kind == 0 ? (int) std::max(1u, blocks_for_host(o.nW))
                                                      : x.kind == 2 ? (int) blocks_for_host(n) : 0));
                bk.push_back(bk.back() + (x.kind != 1 ? rep_check_blocks(o.nlist, n) : 0));
                bm.push_back(bm.back() + (x.kind != 1 ? rep_mem_blocks(n) : 0));
            }
            const int no = (int) ops.size();
            defer_uploads();
            RepOp *dops = upload(batch, ops);
            int *dbr = upload(batch, br), *dbc = upload(batch, bc), *dbk = upload(batch, bk), *dbm = upload(batch, bm);
            flush_uploads();
            HIP_OK(launch_rep_stage(dops, no, dbr, br.back(), dbc, bc.back(), dbk, bk.back(), dbm, bm.back(), entries,
                                    prm.r, prm.min_size, n, keepv, st));
            sync();  // the chunk's scratch and descriptors are reused by the next one
            a = b;
        }
    }
}

// Features of every (tree, dimension) pair of the level (learned and collect
// modes) and, with a pruner loaded, its decisions: a pair whose score is below
// the threshold is skipped, unless its tree has fewer than prune_min_depth
// restricted dimensions.
void Engine::score_pairs(int na) {
    Pool &cur = lvl[cur_lvl];
    lv_pair0.assign(na, 0);
    std::vector<FeatPair> fp;
    std::vector<int> blk{0};
    for (int a = 0; a < na; ++a) {
        const LTree &t = lt[lv_act[a]];
        lv_pair0[a] = (int) fp.size();
        // counts of S are complete for every subspace holding a surviving
        // tree; in phi mode they are decision bits and are left out
        const int *cntS = t.smask && !phi_mode ? subs.at(t.smask).cnt : nullptr;
        const int k = dcount(t.smask);
        for (int m = t.first; m < d; ++m) {
            FeatPair f;
            std::memset(&f, 0, sizeof(f));
            f.W = t.W;
            f.cntS = cntS;
            f.gid = lv_gid + lv_gidoff[a] + (m - t.first) * cells;
            f.mask = t.smask | dbit(m);
            f.nW = t.nW;
            f.pre = t.pre;
            f.m = m;
            f.k = k;
            f.first = t.first;
            f.ngroups = (int) lv_groups[a][m - t.first].size();
            f.bp = nblocks(t.nW);
            f.cap = cap_of(k);
            fp.push_back(f);
            blk.push_back(blk.back() + f.bp);
        }
    }
    const int np = (int) fp.size();
    stats.pairs_scored += np;
    if (np == 0) return;
    const size_t stride = (size_t) kHistBins + cells + 2;
    int *acc = cur.take<int>((size_t) np * stride);
    unsigned long long *accs = cur.take<unsigned long long>(np);
    float *feat = cur.take<float>((size_t) np * kMlpIn);
    HIP_OK(launch_fill_i32(acc, 0, (size_t) np * stride, st));
    HIP_OK(launch_fill_u64(reinterpret_cast<uint64_t *>(accs), 0ull, np, st));
    defer_uploads();
    FeatPair *dfp = upload(cur, fp);
    int *dblk = upload(cur, blk);
    flush_uploads();
    HIP_OK(launch_pair_features(dfp, dblk, np, blk[np], D.Xt, cellC, n, d, cells, prm.min_size, weak_thr, acc, accs,
                                feat, st));
    if (collect) {
        col_base = (long long) col_tree.size();
        const size_t at = col_feat.size();
        col_feat.resize(at + (size_t) np * kMlpIn);
        float *h = reinterpret_cast<float *>(down.get((size_t) np * kMlpIn * sizeof(float)));
        HIP_OK(hipMemcpyAsync(h, feat, (size_t) np * kMlpIn * sizeof(float), hipMemcpyDeviceToHost, st));
        sync();
        std::memcpy(col_feat.data() + at, h, (size_t) np * kMlpIn * sizeof(float));
        for (int a = 0; a < na; ++a)
            for (int m = lt[lv_act[a]].first; m < d; ++m) {
                col_tree.push_back(lv_act[a]);
                col_m.push_back(m);
                col_child0.push_back(-1);
                col_G.push_back(0);
            }
    }
    if (pruner) {
        float *prob = cur.take<float>(np);
        HIP_OK(launch_mlp_predict(feat, np, pruner, prob, st));
        const float *hp = reinterpret_cast<const float *>(down.get((size_t) np * sizeof(float)));
        HIP_OK(hipMemcpyAsync(const_cast<float *>(hp), prob, (size_t) np * sizeof(float), hipMemcpyDeviceToHost, st));
        sync();
        lv_skip.assign(np, 0);
        for (int a = 0; a < na; ++a) {
            const LTree &t = lt[lv_act[a]];
            if (dcount(t.smask) < prune_min_depth) continue;
            for (int m = t.first; m < d; ++m) {
                const int j = lv_pair0[a] + m - t.first;
                if (hp[j] < prune_thr) {
                    lv_skip[j] = 1;
                    stats.pairs_skipped++;
                }
            }
        }
    }
}

// Labels of the collected pairs.  A surviving tree holds a result cluster
// when it was kept by pruneRedundancy and one of its points lies in a cluster
// of the result in its subspace; a pair is positive when a tree below it (a
// child along its dimension, or any descendant of one) holds a result cluster.
void Engine::finish_collect(inscy_result &out) {
    const int nt = (int) lt.size();
    std::vector<TreeHit> th;
    std::vector<int> blk{0};
    for (int i = 1; i < nt; ++i) {
        const LTree &t = lt[i];
        if (!t.pass || t.nW <= 0) continue;
        const Subspace &sb = subs.at(t.smask);
        th.push_back(TreeHit{t.W, sb.comp, sb.inres, t.nW, i});
        blk.push_back(blk.back() + nblocks(t.nW));
    }
    int *hit = store.take<int>(std::max(nt, 1));
    HIP_OK(launch_fill_i32(hit, 0, std::max(nt, 1), st));
    if (!th.empty()) {
        defer_uploads();
        TreeHit *dth = upload(store, th);
        int *dblk = upload(store, blk);
        flush_uploads();
        HIP_OK(launch_tree_hits(dth, dblk, (int) th.size(), blk.back(), hit, st));
    }
    std::vector<int> hh(nt), keep(nt);
    HIP_OK(hipMemcpyAsync(hh.data(), hit, (size_t) nt * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(keep.data(), keepv, (size_t) nt * sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    const long long np = (long long) col_tree.size();
    // pairs of each tree (contiguous in collection order)
    std::vector<long long> p0(nt, -1), pn(nt, 0);
    for (long long j = 0; j < np; ++j) {
        if (p0[col_tree[j]] < 0) p0[col_tree[j]] = j;
        pn[col_tree[j]]++;
    }
    // children have larger indices than their parents
    std::vector<char> contrib(nt, 0);
    std::vector<long long> tcost(nt, 0);
    out.pair_label.assign(np, 0);
    out.pair_cost.assign(2 * np, 0);
    out.pair_info.assign(4 * np, 0);
    for (int i = nt - 1; i >= 0; --i) {
        const LTree &t = lt[i];
        bool c = i > 0 && t.pass && keep[i] && hh[i];
        long long tc = 0;
        for (long long j = p0[i]; j >= 0 && j < p0[i] + pn[i]; ++j) {
            bool pos = false;
            long long cost = t.nW;
            for (int u = 0; u < col_G[j]; ++u) {
                const int ch = col_child0[j] + u;
                pos |= contrib[ch] != 0;
                cost += tcost[ch];
            }
            out.pair_label[j] = pos ? 1 : 0;
            out.pair_cost[2 * j] = t.nW;
            out.pair_cost[2 * j + 1] = cost;
            out.pair_info[4 * j] = dcount(t.smask);
            out.pair_info[4 * j + 1] = col_m[j];
            out.pair_info[4 * j + 2] = t.nW;
            out.pair_info[4 * j + 3] = i;
            c |= pos;
            tc += cost;
        }
        contrib[i] = c;
        tcost[i] = tc;
    }
    out.pair_feat = std::move(col_feat);
    col_feat.clear();
}

void Engine::trace_tree(int f, const std::vector<int> &keep) {
    for (int c : lt[f].kids) {
        const LTree &t = lt[c];
        trace.push_back({(long long) dlo(t.smask), t.lo, t.hi, t.pre, t.nW, t.pass ? 1 : 0,
                         t.pass ? (keep[c] ? 0 : 1) : -1});
        if (t.pass) trace_tree(c, keep);
    }
}

}  // namespace inscy
