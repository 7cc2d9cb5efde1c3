// TEST INFRASTRUCTURE ONLY — the oracle.
//
// A single-threaded CPU restatement of the reference's CPU INSCY path
// (python/inscy.py:121-125 INSCY.CPU -> inscy_map.cpp:230-270 run_cpu3_weak
// -> src/algorithms/inscy/InscyNodeCpu.cu:159-235 INSCYCPU3Weak).  Only the
// tests, the smoke check and the CPU-baseline leg of the benchmark may load
// it, and only as the checker or as the timed CPU baseline; the HIP engine
// never links or calls it.
//
// It keeps the reference's pointer SCY-tree semantics on purpose (maps keyed
// by cell for children and s-connections, points held by leaves, counts of -1
// marking s-connections) because the merge decisions of the algorithm depend
// on exactly which s-connection nodes survive restriction and pruning.  The
// ε-neighbourhood can be produced either by a literal walk of the
// neighbourhood SCY-tree (ScyTreeNode.cu:896-946) or by an index that yields
// the identical ordered candidate list; the tests check the two against each
// other and the index is what makes configurations beyond a few thousand
// points finish.
//
// Build flags: -O2 -ffp-contract=off, no fast-math (the reference's host code
// is x86-64 g++ without FMA contraction).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <set>
#include <string>
#include <unordered_set>
#include <vector>

#include <omp.h>

namespace {

constexpr double kPi = 3.14;  // ClusteringCpu.cu:25 (#define PI 3.14)

// ---------------------------------------------------------------- thresholds
// ClusteringCpu.cu:136-164.  The float/double mix is part of the contract: the
// reference stores double-promoted pow() results into floats.

float gamma_half(int m) {  // Γ(m/2) by the reference's recursion (:136-143)
    if (m == 2) return 1.0f;
    if (m == 1) return (float) std::sqrt(kPi);
    return (float) ((m / 2.0 - 1.0) * (double) gamma_half(m - 2));
}

float unit_ball(int k) {  // c(k) (:145-150)
    float v = (float) std::pow(kPi, k / 2.0);
    v = v / gamma_half(k + 2);
    return v;
}

float alpha_k(int k, float eps, int n) {  // alpha(k) (:152-157)
    float v = (float) ((double) (2 * n) * std::pow((double) eps, (double) k) * (double) unit_ball(k));
    v = (float) ((double) v / (std::pow(1.0, (double) k) * (double) (k + 2)));
    return v;
}

float expden_k(int k, float eps, int n) {  // expDen(k) (:159-164)
    float nc = (float) n * unit_ball(k);
    float v = (float) ((double) nc * std::pow((double) eps, (double) k));
    v = (float) ((double) v / std::pow(1.0, (double) k));
    return v;
}

float omega_k(int k) { return (float) (2.0 / (k + 2.0)); }  // omega(k) (:36-38)

// ---------------------------------------------------------------- data
struct Data {
    const float *X;
    int n, d;
    const float *row(int i) const { return X + (size_t) i * d; }
};

// Euclidean distance in subspace S, accumulated in float in S order with
// separate multiply and add (ClusteringCpu.cu:40-50).
float dist_s(const Data &D, int p, int q, const std::vector<int> &S) {
    const float *a = D.row(p), *b = D.row(q);
    float acc = 0.0f;
    for (int k : S) {
        float diff = a[k] - b[k];
        float sq = diff * diff;
        acc = acc + sq;
    }
    return std::sqrt(acc);
}

// ---------------------------------------------------------------- SCY-tree
// Node fields mirror Node.h:15-60.  count == -1 marks an s-connection.
struct Node {
    int cell;
    int count;
    bool leaf;
    std::map<int, int> kids;   // cell -> node index (regular children)
    std::map<int, int> sconn;  // cell -> node index (s-connections)
    std::vector<int> pts;      // point ids held by a leaf, insertion order
};

struct Tree {
    std::vector<Node> nodes;  // nodes[0] is the root
    std::vector<int> dims;    // level -> data dimension
    std::vector<int> rdims;   // restricted dimensions, in restriction order
    int cells = 0;
    float cell_size = 0.f;
    int npoints = 0;
    bool s_connected = false;

    int add(int cell, int count, bool leaf) {
        Node nd;
        nd.cell = cell;
        nd.count = count;
        nd.leaf = leaf;
        nodes.push_back(std::move(nd));
        return (int) nodes.size() - 1;
    }
    int cell_of(float x) const { return std::min((int) (x / cell_size), cells - 1); }  // ScyTreeNode.cu:100-102
};

int child_or_new(Tree &t, int parent, int cell, bool sconn) {
    std::map<int, int> &m = sconn ? t.nodes[parent].sconn : t.nodes[parent].kids;
    auto it = m.find(cell);
    if (it != m.end()) return it->second;
    int id = t.add(cell, sconn ? -1 : 0, false);
    (sconn ? t.nodes[parent].sconn : t.nodes[parent].kids).emplace(cell, id);
    return id;
}

// ScyTreeNode ctor (ScyTreeNode.cu:105-143) with the s-connection rule of
// construct_s_connection (:34-49).
Tree build_tree(const Data &D, const std::vector<int> &order, int cells, float eps) {
    Tree t;
    t.cells = cells;
    t.cell_size = 1.0f / (float) cells;
    t.dims = order;
    t.add(-1, 0, false);
    const int L = (int) order.size();
    for (int i = 0; i < D.n; ++i) {
        const float *x = D.row(i);
        t.nodes[0].count += 1;
        int cur = 0;
        for (int j = 0; j < L; ++j) {
            float xj = x[order[j]];
            int cj = t.cell_of(xj);
            int child = child_or_new(t, cur, cj, false);
            t.nodes[child].count += 1;
            float upper = (float) (cj + 1) * t.cell_size;
            float bound = upper - eps;
            if (xj >= bound) {
                int s = child_or_new(t, cur, cj, true);
                for (int k = j + 1; k < L; ++k) s = child_or_new(t, s, t.cell_of(x[order[k]]), true);
                t.nodes[s].leaf = true;
            }
            cur = child;
        }
        t.nodes[cur].pts.push_back(i);
        t.nodes[cur].leaf = true;
    }
    t.npoints = t.nodes[0].count;
    return t;
}

// restrict / restrict_node (ScyTreeNode.cu:227-379)
struct Restricter {
    const Tree &T;
    Tree &R;
    int dim_no, cell_no;
    bool flag = false;

    bool visit(int old, int newpar, int depth) {
        const Node &o = T.nodes[old];
        if (T.dims[depth] == dim_no) {
            if (o.cell != cell_no) return false;
            if (R.nodes[newpar].count != -1 && o.count == -1) {
                flag = true;  // s-connection on the restricted dimension under a regular node
                return false;
            }
            if (o.leaf) {
                R.nodes[newpar].leaf = true;
                if (o.count > 0) {
                    R.nodes[newpar].pts.insert(R.nodes[newpar].pts.end(), o.pts.begin(), o.pts.end());
                    R.nodes[newpar].count += o.count;
                }
                return true;
            }
            for (const auto &kv : o.kids) visit(kv.second, newpar, depth + 1);
            for (const auto &kv : o.sconn) visit(kv.second, newpar, depth + 1);
            return true;
        }
        int nn = R.add(o.cell, o.count, o.leaf);
        R.nodes[nn].pts = o.pts;
        bool inc = T.dims[depth] > dim_no;
        if (o.count == -1) {
            for (const auto &kv : o.sconn) {
                bool r = visit(kv.second, nn, depth + 1);
                inc = r || inc;
            }
            if (inc) R.nodes[newpar].sconn.emplace(o.cell, nn);
        } else {
            if (!o.leaf)
                R.nodes[nn].count = 0;
            else
                inc = true;
            for (const auto &kv : o.kids) {
                bool r = visit(kv.second, nn, depth + 1);
                inc = r || inc;
            }
            for (const auto &kv : o.sconn) {
                bool r = visit(kv.second, nn, depth + 1);
                inc = r || inc;
            }
            if (inc) {
                R.nodes[newpar].kids.emplace(o.cell, nn);
                R.nodes[newpar].count += R.nodes[nn].count;
            }
        }
        return inc;
    }
};

Tree restrict_tree(const Tree &T, int dim_no, int cell_no) {
    Tree R;
    R.cells = T.cells;
    R.cell_size = T.cell_size;
    for (int x : T.dims)
        if (x != dim_no) R.dims.push_back(x);
    R.rdims = T.rdims;
    R.rdims.push_back(dim_no);
    R.add(-1, 0, false);
    Restricter rs{T, R, dim_no, cell_no};
    for (const auto &kv : T.nodes[0].kids) rs.visit(kv.second, 0, 0);
    for (const auto &kv : T.nodes[0].sconn) rs.visit(kv.second, 0, 0);
    R.s_connected = rs.flag;
    R.npoints = R.nodes[0].count;
    return R;
}

// Deep-copies the subtree rooted at b of tree B into tree A; returns the new index.
int copy_subtree(Tree &A, const Tree &B, int b) {
    const Node &src = B.nodes[b];
    int a = A.add(src.cell, src.count, src.leaf);
    A.nodes[a].pts = src.pts;
    for (const auto &kv : src.kids) {
        int c = copy_subtree(A, B, kv.second);
        A.nodes[a].kids.emplace(kv.first, c);
    }
    for (const auto &kv : src.sconn) {
        int c = copy_subtree(A, B, kv.second);
        A.nodes[a].sconn.emplace(kv.first, c);
    }
    return a;
}

// mergeNodes (ScyTreeNode.cu:803-831)
void merge_nodes(Tree &A, int a, const Tree &B, int b) {
    if (A.nodes[a].count > 0) {
        A.nodes[a].count += B.nodes[b].count;
        A.nodes[a].pts.insert(A.nodes[a].pts.end(), B.nodes[b].pts.begin(), B.nodes[b].pts.end());
    }
    for (const auto &kv : B.nodes[b].kids) {
        auto it = A.nodes[a].kids.find(kv.first);
        if (it != A.nodes[a].kids.end()) {
            merge_nodes(A, it->second, B, kv.second);
        } else {
            int c = copy_subtree(A, B, kv.second);
            A.nodes[a].kids.emplace(kv.first, c);
        }
    }
    for (const auto &kv : B.nodes[b].sconn) {
        auto it = A.nodes[a].sconn.find(kv.first);
        if (it != A.nodes[a].sconn.end()) {
            merge_nodes(A, it->second, B, kv.second);
        } else {
            int c = copy_subtree(A, B, kv.second);
            A.nodes[a].sconn.emplace(kv.first, c);
        }
    }
}

// mergeWithNeighbors (ScyTreeNode.cu:837-851).  The loop condition reads the
// flag of the most recently restricted neighbour.
void merge_with_neighbors(Tree &R, const Tree &T, int dim_no, int &cell_no) {
    if (!R.s_connected) return;
    bool flag = R.s_connected;
    while (flag && cell_no < R.cells - 1) {
        Tree nb = restrict_tree(T, dim_no, cell_no + 1);
        merge_nodes(R, 0, nb, 0);
        flag = nb.s_connected;
        cell_no++;
    }
    R.npoints = R.nodes[0].count;
    R.s_connected = false;
}

void collect_leaves(const Tree &t, int v, std::vector<int> &out) {  // get_leafs (:422-431)
    const Node &nd = t.nodes[v];
    if (nd.kids.empty() && nd.sconn.empty()) {
        out.push_back(v);
        return;
    }
    for (const auto &kv : nd.kids) collect_leaves(t, kv.second, out);
}

void propagate_counts(Tree &t, int v) {  // propergate_count2 (:446-467)
    if (t.nodes[v].kids.empty() && t.nodes[v].sconn.empty()) return;
    t.nodes[v].count = 0;
    std::vector<int> gone;
    std::vector<std::pair<int, int>> kids(t.nodes[v].kids.begin(), t.nodes[v].kids.end());
    for (const auto &kv : kids) {
        propagate_counts(t, kv.second);
        if (t.nodes[kv.second].count == 0)
            gone.push_back(kv.first);
        else
            t.nodes[v].count += t.nodes[kv.second].count;
    }
    for (int c : gone) {
        t.nodes[v].kids.erase(c);
        t.nodes[v].sconn.erase(c);
    }
}

void collect_points(const Tree &t, int v, std::vector<int> &out) {  // get_points_node (:744-752)
    const Node &nd = t.nodes[v];
    if (nd.kids.empty()) out.insert(out.end(), nd.pts.begin(), nd.pts.end());
    for (const auto &kv : nd.kids) collect_points(t, kv.second, out);
}

// ---------------------------------------------------------------- neighbourhood
struct NeighbourIndex {
    const Data &D;
    const Tree &etree;          // the neighbourhood SCY-tree (ceil(1/eps) cells)
    std::vector<int> order;     // level -> dimension (same order as the SCY-tree)
    bool literal;               // true: walk the tree; false: index
    // index state
    std::vector<int> ecell;     // [n*d] eps-grid cell of point p in dim k
    std::vector<int> rank;      // position of p in the neighbourhood tree's leaf order
    std::vector<std::vector<std::vector<int>>> bucket;  // [dim][cell] -> points

    NeighbourIndex(const Data &D_, const Tree &et, const std::vector<int> &ord, bool lit)
        : D(D_), etree(et), order(ord), literal(lit) {
        if (literal) return;
        const int n = D.n, d = D.d;
        ecell.resize((size_t) n * d);
        for (int p = 0; p < n; ++p)
            for (int k = 0; k < d; ++k) ecell[(size_t) p * d + k] = etree.cell_of(D.row(p)[k]);
        std::vector<int> idx(n);
        for (int i = 0; i < n; ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
            for (int t = 0; t < d; ++t) {
                int ca = ecell[(size_t) a * d + order[t]], cb = ecell[(size_t) b * d + order[t]];
                if (ca != cb) return ca < cb;
            }
            return false;
        });
        rank.resize(n);
        for (int i = 0; i < n; ++i) rank[idx[i]] = i;
        bucket.assign(d, std::vector<std::vector<int>>(etree.cells));
        for (int p = 0; p < n; ++p)
            for (int k = 0; k < d; ++k) bucket[k][ecell[(size_t) p * d + k]].push_back(p);
    }

    // Dimensions of S that the tree walk filters on: a greedy in-order match of
    // S against the level order (ScyTreeNode.cu:910-915).
    std::vector<int> matched(const std::vector<int> &S) const {
        std::vector<int> m;
        size_t si = 0;
        for (int t = 0; t < (int) order.size() && si < S.size(); ++t)
            if (order[t] == S[si]) m.push_back(S[si++]);
        return m;
    }

    void walk(std::vector<int> &list, const float *p, int v, int depth, size_t si, const std::vector<int> &S) const {
        const Node &nd = etree.nodes[v];
        if (nd.kids.empty()) {
            list.insert(list.end(), nd.pts.begin(), nd.pts.end());
            return;
        }
        depth += 1;
        bool filt = si < S.size() && order[depth] == S[si];
        int center = 0;
        if (filt) {
            center = etree.cell_of(p[S[si]]);
            si += 1;
        }
        for (const auto &kv : nd.kids) {
            if (!filt || (center - 1 <= kv.first && kv.first <= center + 1)) walk(list, p, kv.second, depth, si, S);
        }
    }

    // The candidate set is the points within one ε-cell of p in every matched
    // dimension; it is drawn from the bucket of the matched dimension with the
    // fewest points in p's three cells, then checked on the others.  ordered:
    // the neighbourhood tree's leaf order (needed where the list's order
    // matters: φ's float sum); otherwise the bucket order.
    std::vector<int> candidates(int p, const std::vector<int> &S, bool ordered = true) const {
        std::vector<int> list;
        if (literal) {
            walk(list, D.row(p), 0, -1, 0, S);
            return list;
        }
        std::vector<int> M = matched(S);
        const int d = D.d;
        size_t best = 0, best_size = SIZE_MAX;
        for (size_t t = 0; t < M.size(); ++t) {
            int ct = ecell[(size_t) p * d + M[t]];
            size_t sz = 0;
            for (int c = std::max(0, ct - 1); c <= std::min(etree.cells - 1, ct + 1); ++c) sz += bucket[M[t]][c].size();
            if (sz < best_size) best_size = sz, best = t;
        }
        int m0 = M[best];
        int c0 = ecell[(size_t) p * d + m0];
        for (int c = std::max(0, c0 - 1); c <= std::min(etree.cells - 1, c0 + 1); ++c) {
            for (int q : bucket[m0][c]) {
                bool ok = true;
                for (size_t t = 0; t < M.size() && ok; ++t) {
                    if (t == best) continue;
                    int dq = ecell[(size_t) q * d + M[t]] - ecell[(size_t) p * d + M[t]];
                    ok = dq >= -1 && dq <= 1;
                }
                if (ok) list.push_back(q);
            }
        }
        if (ordered) std::sort(list.begin(), list.end(), [&](int a, int b) { return rank[a] < rank[b]; });
        return list;
    }

    // neighborhood() (ClusteringCpu.cu:52-81): ε-neighbours of p in S, in the
    // neighbourhood tree's leaf order when ordered.
    std::vector<int> neighbours(int p, const std::vector<int> &S, float eps, bool ordered = true) const {
        std::vector<int> out;
        for (int q : candidates(p, S, ordered)) {
            if (q == p) continue;
            float dd = dist_s(D, p, q, S);
            if (eps >= dd) out.push_back(q);
        }
        return out;
    }
};

// phi() (ClusteringCpu.cu:116-124): float accumulator, double term.
float phi(const Data &D, int p, const std::vector<int> &nb, float eps, const std::vector<int> &S) {
    float sum = 0.0f;
    for (int q : nb) {
        float r = dist_s(D, p, q, S) / eps;
        float sq = r * r;
        sum = (float) ((double) sum + (1.0 - (double) sq));
    }
    return sum;
}

// ---------------------------------------------------------------- result map
// vec_cmp (util.cu:633-653): lexicographic from the last element backwards.
struct VecCmp {
    bool operator()(const std::vector<int> &a, const std::vector<int> &b) const {
        int i = (int) a.size() - 1, j = (int) b.size() - 1;
        while (a[i] == b[j]) {
            --i;
            --j;
            if (i < 0 || j < 0) return i < j;
        }
        return a[i] < b[j];
    }
};
using Result = std::map<std::vector<int>, std::vector<int>, VecCmp>;

bool is_subspace_of(const std::vector<int> &S, const std::vector<int> &H) {  // subspace_of (util.cu:619-631)
    size_t i = 0, j = 0;
    while (j < H.size() && i < S.size()) {
        if (S[i] == H[j]) ++i;
        ++j;
    }
    return i == S.size();
}

int vmax(const std::vector<int> &v) {  // v_max (util.cu:568-576)
    int m = -100000;
    for (int x : v)
        if (x > m) m = x;
    return m;
}

// join (util.cu:655-743)
void join(Result &result, std::vector<int> &cl, const std::vector<int> &S, int min_size, float r) {
    if (vmax(cl) < 0) return;
    const int n = (int) cl.size();
    std::map<int, int> sizes;
    for (int x : cl)
        if (x >= 0) sizes[x]++;
    for (int i = 0; i < n; ++i)
        if (cl[i] >= 0 && sizes[cl[i]] < min_size) cl[i] = -1;
    for (const auto &entry : result) {
        if (!is_subspace_of(S, entry.first)) continue;
        const std::vector<int> &H = entry.second;
        std::map<int, int> sizes_h;
        for (int x : H)
            if (x >= 0) sizes_h[x]++;
        std::set<int> drop;
        for (int i = 0; i < n; ++i) {
            int a = cl[i], b = H[i];
            if (a >= 0 && b >= 0 && (float) sizes[a] * r < (float) sizes_h[b]) drop.insert(a);
        }
        for (int i = 0; i < n; ++i)
            if (cl[i] >= 0 && drop.count(cl[i])) cl[i] = -1;
    }
    if (vmax(cl) < 0) return;
    auto it = result.find(S);
    if (it == result.end()) {
        result.emplace(S, cl);
        return;
    }
    std::vector<int> &old = it->second;
    int m = vmax(old);
    for (int i = 0; i < n; ++i) {
        if (cl[i] == -2)
            old[i] = -2;
        else if (cl[i] >= 0)
            old[i] = m + 1 + cl[i];
    }
}

// ---------------------------------------------------------------- driver
struct TraceRow {
    uint64_t mask;
    int lo, hi;       // merged cell range on the restricted dimension
    int pre, post;    // points before / after the weak-density prune
    int pass;         // prune outcome
    int redundant;    // -1: not evaluated; 0: kept; 1: pruned as redundant
};

struct Ctx {
    const Data &D;
    const NeighbourIndex &nbr;
    float eps, F, r;
    int num_obj, min_size;
    bool rect;
    Result result;
    long long calls = 0;
    long long trees = 0;        // restricted trees evaluated by the prune step
    long long queries = 0;      // ε-neighbourhood queries issued
    long long nbr_total = 0;    // Σ |N_S(p)| over all queries
    bool trace_on = false;
    std::vector<TraceRow> trace;
    // thresholds
    float weak_rect_thr, weak_phi_thr;
    // optional wall-clock budget (CPU-baseline sampling): the run stops at the
    // first prune after the deadline
    bool timed = false;
    std::chrono::steady_clock::time_point deadline;
};

struct Deadline {};

// the set S as a bitmask (d <= 128); trace rows keep its low 64 bits
unsigned __int128 mask_of(const std::vector<int> &S) {
    unsigned __int128 m = 0;
    for (int k : S) m |= (unsigned __int128) 1 << k;
    return m;
}

// pruneRecursionAndRemove2 (ScyTreeNode.cu:582-667)
bool prune(Tree &R, Ctx &C) {
    std::vector<int> leaves;
    collect_leaves(R, 0, leaves);
    const std::vector<int> &S = R.rdims;
    int kept = 0;
    for (int lv : leaves) {
        std::vector<int> keep;
        int cnt = 0;
        for (int p : R.nodes[lv].pts) {
            std::vector<int> nb = C.nbr.neighbours(p, S, C.eps);
            C.queries++;
            C.nbr_total += (long long) nb.size();
            bool dense;
            if (C.rect)
                dense = (float) nb.size() >= C.weak_rect_thr;
            else
                dense = phi(C.D, p, nb, C.eps, S) >= C.weak_phi_thr;
            if (dense) {
                ++kept;
                ++cnt;
                keep.push_back(p);
            }
        }
        R.nodes[lv].pts = keep;
        R.nodes[lv].count = cnt;
    }
    propagate_counts(R, 0);
    R.npoints = R.nodes[0].count;
    return kept >= C.min_size;
}

// pruneRedundancy (ScyTreeNode.cu:669-734)
bool prune_redundancy(const Tree &R, Ctx &C) {
    int max_min = 0;
    std::vector<int> pts;
    collect_points(R, 0, pts);
    for (const auto &entry : C.result) {
        if (!is_subspace_of(R.rdims, entry.first)) continue;
        const std::vector<int> &H = entry.second;
        std::map<int, int> size;
        std::map<int, bool> used;
        for (int x : H)
            if (x >= 0) {
                size[x]++;
                used.emplace(x, false);
            }
        for (int p : pts)
            if (H[p] >= 0) used[H[p]] = true;
        int mn = -1;
        for (const auto &kv : size)
            if (used[kv.first] && (mn == -1 || kv.second < mn)) mn = kv.second;
        if (mn > max_min) max_min = mn;
    }
    if (max_min == 0) return true;
    float lhs = (float) R.npoints * C.r;
    return (double) lhs > (double) max_min * 1.0;
}

// INSCYClusteringImplCPUAll (ClusteringCpu.cu:317-389)
void cluster(const Tree &R, Ctx &C, std::vector<int> &cl) {
    const std::vector<int> &S = R.rdims;
    const int k = (int) S.size();
    const int n = C.D.n;
    int label = std::max(0, vmax(cl)) + 1;
    std::vector<int> pts;
    collect_points(R, 0, pts);
    float thr_rect = std::max(C.F * expden_k(k, C.eps, n), (float) C.num_obj);
    float thr_phi = std::max(C.F * alpha_k(k, C.eps, n), (float) C.num_obj * omega_k(k));
    int labelled = 0, before = 0;
    std::queue<int> q;
    for (int s : pts) {
        if (cl[s] != -1) continue;
        before = labelled;
        q.push(s);
        while (!q.empty()) {
            int p = q.front();
            q.pop();
            std::vector<int> nb = C.nbr.neighbours(p, S, C.eps);
            C.queries++;
            C.nbr_total += (long long) nb.size();
            bool dense = C.rect ? (float) nb.size() >= thr_rect : phi(C.D, p, nb, C.eps, S) >= thr_phi;
            if (!dense) continue;
            cl[p] = label;
            ++labelled;
            for (int x : nb)
                if (cl[x] == -1) {
                    cl[x] = -2;
                    q.push(x);
                }
        }
        if (labelled > before) ++label;
    }
}

void inscy(const Tree &T, Ctx &C, int first) {
    C.calls++;
    const int d = C.D.d;
    for (int dim = first; dim < d; ++dim) {
        std::vector<int> cl(C.D.n, -1);
        std::vector<int> S;
        for (int cell = 0; cell < T.cells; ++cell) {
            Tree R = restrict_tree(T, dim, cell);
            S = R.rdims;
            int lo = cell;
            merge_with_neighbors(R, T, dim, cell);
            std::vector<int> pre_pts;
            collect_points(R, 0, pre_pts);
            int pre = (int) pre_pts.size();  // points of the merged region before the prune
            if (C.timed && std::chrono::steady_clock::now() > C.deadline) throw Deadline{};
            bool pass = prune(R, C);
            C.trees++;
            TraceRow row{(uint64_t) mask_of(S), lo, cell, pre, R.npoints, pass ? 1 : 0, -1};
            size_t trace_at = C.trace.size();
            if (C.trace_on) C.trace.push_back(row);
            if (pass) {
                inscy(R, C, dim + 1);
                bool keep = prune_redundancy(R, C);
                if (C.trace_on) C.trace[trace_at].redundant = keep ? 0 : 1;
                if (keep) cluster(R, C, cl);
            }
        }
        join(C.result, cl, S, C.min_size, C.r);
    }
}

// get_subspace_order (inscy_map.cpp:61-121)
std::vector<int> dimension_order(const Data &D, int cells, int entropy_order) {
    const int n = D.n, d = D.d;
    std::vector<int> order(d);
    for (int i = 0; i < d; ++i) order[i] = i;
    if (entropy_order == 0) return order;
    std::vector<std::vector<int>> count(d, std::vector<int>(cells, 0));
    float cs = 1.0f / (float) cells;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < d; ++j) count[j][std::min((int) (D.row(i)[j] / cs), cells - 1)] += 1;
    std::vector<std::pair<float, int>> ent(d);
    for (int j = 0; j < d; ++j) {
        float e = 0.0f;
        for (int c = 0; c < cells; ++c) {
            float p = (float) count[j][c] / (float) n;
            e -= p * std::log(p);
        }
        ent[j] = {e, j};
    }
    std::sort(ent.begin(), ent.end());
    for (int j = 0; j < d; ++j) order[j] = ent[j].second;
    if (entropy_order == -1) std::reverse(order.begin(), order.end());
    return order;
}

struct Run {
    std::vector<std::vector<int>> subspaces;
    std::vector<std::vector<int>> labels;
    std::vector<TraceRow> trace;
    long long calls = 0, trees = 0, queries = 0, nbr_total = 0;
};


// ---------------------------------------------------------------- set-based restatement
// The same algorithm with the SCY-tree expressed as sets instead of pointers:
// a tree is (restricted dims S, point set P, s-connection paths SP), where an
// s-connection path (b, x) is the chain that point x starts at the level of
// dimension b because x lies within ε of the upper edge of its cell there
// (ScyTreeNode.cu:34-49).  With the identity dimension order:
//  * restrict+merge on dim m with cell group G keeps points and paths whose
//    cell in m is in G; paths branching at m are consumed by the merge test
//    (ScyTreeNode.cu:262-267 sets is_s_connected) and paths branching below m
//    can never matter again;
//  * the weak-density prune keeps W ⊆ P; a path (b, x) survives iff the
//    regular node it hangs from keeps a surviving point (or is the root) and
//    the same-cell regular sibling it shadows is not one that the prune
//    emptied (propergate_count2, ScyTreeNode.cu:446-467);
//  * clustering labels every dense component (dense-dense ε-graph in S over
//    the whole data) that holds a dense point of a kept region, which is what
//    the reference's BFS (ClusteringCpu.cu:317-389) produces up to label names.
// Densities are memoised per subspace and computed with all host threads.

struct SPath {
    int b;
    int x;
};

struct STree {
    std::vector<int> S;
    std::vector<int> pts;
    std::vector<SPath> sp;
    std::vector<struct Chain> ch;  // permuted level order: the s-connection chains
};

// An s-connection chain of a permuted-level-order tree: it starts at the level
// of dimension b under point x's regular path and runs through x's cells down
// to level `end` (SCY-tree level numbering; restricted levels are skipped).
struct Chain {
    int b;
    int x;
    int end;
};

struct U128Hash {
    size_t operator()(const std::pair<unsigned __int128, int> &k) const {
        uint64_t lo = (uint64_t) k.first, hi = (uint64_t) (k.first >> 64);
        uint64_t h = lo * 0x9E3779B97F4A7C15ull ^ (hi + 0x632BE59BD9B4E019ull + (lo << 6) + (lo >> 2));
        h ^= (uint64_t) k.second * 0xC2B2AE3D27D4EB4Full;
        h ^= h >> 29;
        return (size_t) h;
    }
};
using PrefixSet = std::unordered_set<std::pair<unsigned __int128, int>, U128Hash>;

struct SubspaceMemo {
    std::vector<int> cnt;        // |N_S(p)|, -1 = not computed
    std::vector<float> phiv;     // phi_S(p) (non-rectangular mode)
    bool comps = false;
    std::vector<int> comp;       // component root (min point id) of dense p, else -1
    std::vector<int> compsize;   // indexed by root id
};

struct SetEngine {
    const Data &D;
    const NeighbourIndex &nbr;
    float eps, F, r;
    int num_obj, min_size, cells;
    bool rect;
    std::vector<uint8_t> cellc;  // [p*d+k]
    std::map<unsigned __int128, SubspaceMemo> memo;
    std::map<std::vector<int>, std::set<int>, VecCmp> result;
    long long calls = 0, trees = 0, queries = 0;
    bool trace_on = false;
    std::vector<TraceRow> trace;
    float weak_rect_thr, weak_phi_thr;

    SetEngine(const Data &D_, const NeighbourIndex &nb, float eps_, float F_, float r_, int mu, int ms, int c,
              bool rect_)
        : D(D_), nbr(nb), eps(eps_), F(F_), r(r_), num_obj(mu), min_size(ms), cells(c), rect(rect_) {
        float cs = 1.0f / (float) cells;
        cellc.resize((size_t) D.n * D.d);
        for (int p = 0; p < D.n; ++p)
            for (int k = 0; k < D.d; ++k)
                cellc[(size_t) p * D.d + k] = (uint8_t) std::min((int) (D.row(p)[k] / cs), cells - 1);
        weak_rect_thr = std::max(F * expden_k(D.d, eps, D.n), (float) num_obj);
        weak_phi_thr = std::max(F * alpha_k(D.d, eps, D.n), (float) num_obj * omega_k(D.d));
    }

    int cc(int p, int k) const { return cellc[(size_t) p * D.d + k]; }

    SubspaceMemo &memo_of(const std::vector<int> &S) {
        SubspaceMemo &m = memo[mask_of(S)];
        if (m.cnt.empty()) {
            m.cnt.assign(D.n, -1);
            if (!rect) m.phiv.assign(D.n, 0.f);
        }
        return m;
    }

    void ensure(const std::vector<int> &S, const std::vector<int> &pts) {
        SubspaceMemo &m = memo_of(S);
        std::vector<int> todo;
        for (int p : pts)
            if (m.cnt[p] < 0) todo.push_back(p);
        queries += (long long) todo.size();
#pragma omp parallel for schedule(dynamic, 64)
        for (long i = 0; i < (long) todo.size(); ++i) {
            int p = todo[i];
            std::vector<int> nb = nbr.neighbours(p, S, eps, !rect);  // only φ's sum depends on the order
            if (!rect) m.phiv[p] = phi(D, p, nb, eps, S);
            m.cnt[p] = (int) nb.size();
        }
    }

    bool weak(const SubspaceMemo &m, int p) const {
        return rect ? (float) m.cnt[p] >= weak_rect_thr : m.phiv[p] >= weak_phi_thr;
    }

    bool dense(const SubspaceMemo &m, int p, int k) const {
        if (rect) return (float) m.cnt[p] >= std::max(F * expden_k(k, eps, D.n), (float) num_obj);
        return m.phiv[p] >= std::max(F * alpha_k(k, eps, D.n), (float) num_obj * omega_k(k));
    }

    static int find(std::vector<int> &par, int a) {
        while (par[a] != a) {
            par[a] = par[par[a]];
            a = par[a];
        }
        return a;
    }

    SubspaceMemo &components(const std::vector<int> &S) {
        SubspaceMemo &m = memo_of(S);
        if (m.comps) return m;
        std::vector<int> all(D.n);
        for (int i = 0; i < D.n; ++i) all[i] = i;
        ensure(S, all);
        const int k = (int) S.size();
        std::vector<char> dn(D.n);
        for (int p = 0; p < D.n; ++p) dn[p] = dense(m, p, k);
        std::vector<std::vector<int>> lists(D.n);
#pragma omp parallel for schedule(dynamic, 64)
        for (int p = 0; p < D.n; ++p)
            if (dn[p]) lists[p] = nbr.neighbours(p, S, eps, false);  // components: order-free
        std::vector<int> par(D.n);
        for (int i = 0; i < D.n; ++i) par[i] = i;
        for (int p = 0; p < D.n; ++p) {
            if (!dn[p]) continue;
            for (int q : lists[p]) {
                if (!dn[q]) continue;
                int a = find(par, p), b = find(par, q);
                if (a == b) continue;
                if (a < b)
                    par[b] = a;
                else
                    par[a] = b;
            }
        }
        m.comp.assign(D.n, -1);
        m.compsize.assign(D.n, 0);
        for (int p = 0; p < D.n; ++p)
            if (dn[p]) {
                m.comp[p] = find(par, p);
                m.compsize[m.comp[p]]++;
            }
        m.comps = true;
        return m;
    }

    // prefix keys of point z over the levels of a tree with restricted set S
    void digits(int z, const std::vector<int> &levels, std::vector<uint8_t> &out) const {
        out.resize(levels.size());
        for (size_t t = 0; t < levels.size(); ++t) out[t] = (uint8_t) cc(z, levels[t]);
    }

    void insert_prefixes(PrefixSet &set, int z, const std::vector<int> &levels, int upto) const {
        unsigned __int128 key = 0;
        for (int t = 0; t < upto; ++t) {
            key = key * (unsigned) cells + (unsigned) cc(z, levels[t]);
            set.emplace(key, t + 1);
        }
    }

    unsigned __int128 prefix_key(int z, const std::vector<int> &levels, int t) const {
        unsigned __int128 key = 0;
        for (int i = 0; i < t; ++i) key = key * (unsigned) cells + (unsigned) cc(z, levels[i]);
        return key;
    }

    std::vector<SPath> prune_paths(const STree &R, const std::vector<int> &W) const {
        std::vector<int> levels;
        {
            size_t si = 0;
            for (int k = 0; k < D.d; ++k) {
                if (si < R.S.size() && R.S[si] == k) {
                    ++si;
                    continue;
                }
                levels.push_back(k);
            }
        }
        const int L = (int) levels.size();
        std::vector<int> level_of(D.d, -1);
        for (int t = 0; t < L; ++t) level_of[levels[t]] = t;
        PrefixSet pw, pp, pa;
        for (int w : W) insert_prefixes(pw, w, levels, L);
        for (int p : R.pts) insert_prefixes(pp, p, levels, L);
        for (const SPath &s : R.sp) insert_prefixes(pa, s.x, levels, level_of[s.b]);
        std::vector<SPath> out;
        // mutation knob for checking the sensitivity of the cross-check tests:
        // 1 drops the prune of paths, 2 ignores the parent rule, 3 ignores the
        // sibling rule, 4 ignores anchor prefixes.
        static const int mutate = getenv("ORACLE_SETS_MUTATE") ? atoi(getenv("ORACLE_SETS_MUTATE")) : 0;
        if (mutate == 1) return R.sp;
        for (const SPath &s : R.sp) {
            int t = level_of[s.b];
            bool a = t == 0 || pw.count({prefix_key(s.x, levels, t), t});
            if (mutate == 2 || mutate == 5) a = true;
            if (!a) continue;
            auto q = std::make_pair(prefix_key(s.x, levels, t + 1), t + 1);
            // pa holds the regular-ancestor prefixes (lengths 1..t_b2) of every path, so
            // a hit at length t+1 means a path branching below level t hangs under q.
            bool b_ok = pw.count(q) || (!pp.count(q) && !pa.count(q));
            if (mutate == 3) b_ok = true;
            if (mutate == 5) b_ok = pw.count(q) != 0;  // the simplified rule the engine uses
            if (mutate == 4) b_ok = pw.count(q) || !pp.count(q);
            if (b_ok) out.push_back(s);
        }
        return out;
    }

    // ---- permuted level order (entropy_order != 0)
    // lvl[k]: SCY-tree level of dimension k; ord[t]: dimension at level t.
    std::vector<int> lvl, ord;

    // the unrestricted dimensions of a tree with restricted set S, in level order
    std::vector<int> levels_of(const std::vector<int> &S) const {
        std::vector<char> in(D.d, 0);
        for (int k : S) in[k] = 1;
        std::vector<int> out;
        for (int t = 0; t < D.d; ++t)
            if (!in[ord[t]]) out.push_back(ord[t]);
        return out;
    }

    // restrict_node (ScyTreeNode.cu:227-320) on one chain of the tree T
    // (restricted set S_T) for dimension m and cell group [lo, hi]: whether
    // the chain is in the restricted tree, and its new last level.  A chain
    // node off the restricted level is kept when its dimension id exceeds m
    // or a node below it is kept; the chain node on the restricted level
    // passes (and keeps the chain above it) when its cell is in the group.
    bool restrict_chain(const Chain &c, int m, int lo, int hi, const std::vector<char> &restricted, Chain &out) const {
        const int Lb = lvl[c.b], Lm = lvl[m];
        const int v = cc(c.x, m);
        const bool cell_in = v >= lo && v <= hi;
        auto deepest_above_m = [&](int from, int to) {  // deepest unrestricted level in [from, to] with dim > m
            int best = -1;
            for (int t = from; t <= to; ++t)
                if (t != Lm && !restricted[ord[t]] && ord[t] > m) best = t;
            return best;
        };
        out = c;
        if (Lb > Lm) {
            if (!cell_in) return false;  // hangs below the restricted level, under x's cell
            const int e = deepest_above_m(Lb, c.end);
            if (e < 0) return false;
            out.end = e;
            return true;
        }
        const bool passes = c.end >= Lm && cell_in;
        if (passes) {
            int e = deepest_above_m(Lm + 1, c.end);
            if (e < 0) {
                e = Lm - 1;  // the chain node above the restricted level
                while (e > Lb && restricted[ord[e]]) --e;
            }
            out.end = e;
            return true;
        }
        const int e = deepest_above_m(Lb, std::min(c.end, Lm - 1));
        if (e < 0) return false;
        out.end = e;
        return true;
    }

    // pruneRecursionAndRemove2 + propergate_count2 (ScyTreeNode.cu:446-467,
    // 582-667) on the chains of a restricted tree R (from T by dimension m):
    // a chain stays when its regular parent keeps a weak-dense point, unless
    // its same-cell regular sibling existed before the prune and lost its
    // points.  Regular nodes before the prune: prefixes of R's points, the
    // nodes of T above the restricted level kept because a level between
    // them and the restricted one has a dimension id above m, and the
    // ancestors of R's chains.
    std::vector<Chain> prune_chains(const STree &R, const std::vector<int> &W, const std::vector<int> &Tpts, int m) const {
        const std::vector<int> levels = levels_of(R.S);
        const int L = (int) levels.size();
        std::vector<int> pos(D.d, -1);
        for (int t = 0; t < L; ++t) pos[levels[t]] = t;
        const int Lm = lvl[m];
        // level_cond[t]: the node at child level index t lies above level Lm
        // and some unrestricted level between it and Lm has a dimension > m
        std::vector<char> restricted(D.d, 0);
        for (int k : R.S) restricted[k] = 1;
        std::vector<char> level_cond(L, 0);
        for (int t = 0; t < L; ++t) {
            const int Lt = lvl[levels[t]];
            if (Lt >= Lm) continue;
            for (int u = Lt; u < Lm; ++u)
                if (!restricted[ord[u]] && ord[u] > m) level_cond[t] = 1;
        }
        PrefixSet pw, pp, pt, pa;
        for (int w : W) insert_prefixes(pw, w, levels, L);
        for (int p : R.pts) insert_prefixes(pp, p, levels, L);
        for (int p : Tpts) insert_prefixes(pt, p, levels, L);
        for (const Chain &c : R.ch) insert_prefixes(pa, c.x, levels, pos[c.b]);
        std::vector<Chain> out;
        for (const Chain &c : R.ch) {
            const int t = pos[c.b];
            const bool parent = t == 0 || pw.count({prefix_key(c.x, levels, t), t});
            if (!parent) continue;
            const auto q = std::make_pair(prefix_key(c.x, levels, t + 1), t + 1);
            bool existed = pp.count(q) || pa.count(q) || (level_cond[t] && pt.count(q));
            // mutation knob (see prune above): 7 ignores the sibling rule, 8
            // ignores the empty nodes kept above m's level
            static const int mutate = getenv("ORACLE_SETS_MUTATE") ? atoi(getenv("ORACLE_SETS_MUTATE")) : 0;
            if (mutate == 7) existed = false;
            if (mutate == 8) existed = pp.count(q) || pa.count(q);
            if (pw.count(q) || !existed) out.push_back(c);
        }
        return out;
    }

    // mergeWithNeighbors merges the trees of cells lo+1.. into the tree of
    // cell lo, and mergeNodes (ScyTreeNode.cu:803-833) adds a node's count and
    // points only into a node whose count is positive.  When m's level is the
    // deepest unrestricted one, the points sit one level up, where the tree of
    // cell lo keeps an empty node for every prefix of T when that level's
    // dimension id exceeds m: the points of such a prefix are then kept only
    // if one of them lies in cell lo.  With no level left above, the root
    // plays that node.
    std::vector<int> merged_points(const STree &T, const std::vector<int> &pts, int m, int lo,
                                   const std::vector<char> &restricted) const {
        // mutation knob (see prune above): 6 ignores mergeNodes' positive-count rule
        static const int mutate = getenv("ORACLE_SETS_MUTATE") ? atoi(getenv("ORACLE_SETS_MUTATE")) : 0;
        if (mutate == 6) return pts;
        const int Lm = lvl[m];
        std::vector<int> above;
        for (int t = 0; t < D.d; ++t) {
            if (restricted[ord[t]] || t == Lm) continue;
            if (t > Lm) return pts;  // m's level is not the deepest: nothing is lost
            above.push_back(ord[t]);
        }
        if (!above.empty() && above.back() < m) return pts;  // no empty node at the point level
        std::set<unsigned __int128> with_lo;
        const int A = (int) above.size();
        for (int p : pts)
            if (cc(p, m) == lo) with_lo.insert(prefix_key(p, above, A));
        std::vector<int> out;
        for (int p : pts)
            if (with_lo.count(prefix_key(p, above, A))) out.push_back(p);
        return out;
    }

    void rec_perm(const STree &T, int first) {
        calls++;
        std::vector<char> restricted(D.d, 0);
        for (int k : T.S) restricted[k] = 1;
        for (int m = first; m < D.d; ++m) {
            std::vector<int> S = T.S;
            S.push_back(m);
            std::vector<char> flag(cells, 0);
            for (const Chain &c : T.ch)
                if (c.b == m) flag[cc(c.x, m)] = 1;
            std::set<int> labelled;
            for (int c = 0; c < cells; ++c) {
                int lo = c;
                while (flag[c] && c < cells - 1) ++c;
                int hi = c;
                STree R;
                R.S = S;
                for (int p : T.pts) {
                    int v = cc(p, m);
                    if (v >= lo && v <= hi) R.pts.push_back(p);
                }
                for (const Chain &ch : T.ch) {
                    if (ch.b == m) continue;
                    Chain o;
                    if (restrict_chain(ch, m, lo, hi, restricted, o)) R.ch.push_back(o);
                }
                if (hi > lo) R.pts = merged_points(T, R.pts, m, lo, restricted);
                ensure(S, R.pts);
                SubspaceMemo &ms = memo_of(S);
                std::vector<int> W;
                for (int p : R.pts)
                    if (weak(ms, p)) W.push_back(p);
                trees++;
                bool pass = (int) W.size() >= min_size;
                size_t at = trace.size();
                if (trace_on)
                    trace.push_back(TraceRow{(uint64_t) mask_of(S), lo, hi, (int) R.pts.size(), (int) W.size(), pass ? 1 : 0, -1});
                if (!pass) continue;
                R.ch = prune_chains(R, W, T.pts, m);
                R.pts = W;
                rec_perm(R, m + 1);
                bool keep = prune_redundancy(R, W);
                if (trace_on) trace[at].redundant = keep ? 0 : 1;
                if (!keep) continue;
                SubspaceMemo &mc = components(S);
                for (int w : W)
                    if (mc.comp[w] >= 0) labelled.insert(mc.comp[w]);
            }
            join(S, labelled);
        }
    }

    bool prune_redundancy(const STree &R, const std::vector<int> &W) {
        int max_min = 0;
        for (auto &entry : result) {
            if (!is_subspace_of(R.S, entry.first)) continue;
            SubspaceMemo &mh = components(entry.first);
            int mn = -1;
            for (int w : W) {
                int k = mh.comp[w];
                if (k < 0 || !entry.second.count(k)) continue;
                int sz = mh.compsize[k];
                if (mn == -1 || sz < mn) mn = sz;
            }
            if (mn > max_min) max_min = mn;
        }
        if (max_min == 0) return true;
        float lhs = (float) (int) W.size() * r;
        return (double) lhs > (double) max_min;
    }

    void join(const std::vector<int> &S, const std::set<int> &labelled) {
        if (labelled.empty()) return;
        SubspaceMemo &ms = components(S);
        std::set<int> keep;
        for (int k : labelled)
            if (ms.compsize[k] >= min_size) keep.insert(k);
        for (auto &entry : result) {
            if (keep.empty()) break;
            if (!is_subspace_of(S, entry.first)) continue;
            SubspaceMemo &mh = components(entry.first);
            std::set<int> drop;
            for (int i = 0; i < D.n; ++i) {
                int k = ms.comp[i];
                if (k < 0 || !keep.count(k)) continue;
                int kh = mh.comp[i];
                if (kh < 0 || !entry.second.count(kh)) continue;
                if ((float) ms.compsize[k] * r < (float) mh.compsize[kh]) drop.insert(k);
            }
            for (int k : drop) keep.erase(k);
        }
        if (keep.empty()) return;
        result[S].insert(keep.begin(), keep.end());
    }

    void rec(const STree &T, int first) {
        calls++;
        for (int m = first; m < D.d; ++m) {
            std::vector<int> S = T.S;
            S.push_back(m);
            std::vector<char> flag(cells, 0);
            for (const SPath &s : T.sp)
                if (s.b == m) flag[cc(s.x, m)] = 1;
            std::set<int> labelled;
            for (int c = 0; c < cells; ++c) {
                int lo = c;
                while (flag[c] && c < cells - 1) ++c;
                int hi = c;
                STree R;
                R.S = S;
                for (int p : T.pts) {
                    int v = cc(p, m);
                    if (v >= lo && v <= hi) R.pts.push_back(p);
                }
                for (const SPath &s : T.sp)
                    if (s.b > m) {
                        int v = cc(s.x, m);
                        if (v >= lo && v <= hi) R.sp.push_back(s);
                    }
                ensure(S, R.pts);
                SubspaceMemo &ms = memo_of(S);
                std::vector<int> W;
                for (int p : R.pts)
                    if (weak(ms, p)) W.push_back(p);
                trees++;
                bool pass = (int) W.size() >= min_size;
                size_t at = trace.size();
                if (trace_on)
                    trace.push_back(TraceRow{(uint64_t) mask_of(S), lo, hi, (int) R.pts.size(), (int) W.size(), pass ? 1 : 0, -1});
                if (!pass) continue;
                R.sp = prune_paths(R, W);
                R.pts = W;
                rec(R, m + 1);
                bool keep = prune_redundancy(R, W);
                if (trace_on) trace[at].redundant = keep ? 0 : 1;
                if (!keep) continue;
                SubspaceMemo &mc = components(S);
                for (int w : W)
                    if (mc.comp[w] >= 0) labelled.insert(mc.comp[w]);
            }
            join(S, labelled);
        }
    }
};

}  // namespace

// ---------------------------------------------------------------- C ABI (test use)
extern "C" {

// Runs the oracle.  nbr_literal=1 walks the neighbourhood tree literally.
void *oracle_ref_run(const float *X, int n, int d, float eps, float F, int num_obj, int min_size, float r,
                     int cells, int rectangular, int entropy_order, int nbr_literal, int trace) {
    Data D{X, n, d};
    std::vector<int> order = dimension_order(D, cells, entropy_order);
    Tree root = build_tree(D, order, cells, eps);
    int ecells = (int) std::ceil(1.0 / (double) eps);
    Tree etree = build_tree(D, order, ecells, eps);
    NeighbourIndex nbr(D, etree, order, nbr_literal != 0);
    Ctx C{D, nbr, eps, F, r, num_obj, min_size, rectangular != 0};
    C.trace_on = trace != 0;
    C.weak_rect_thr = std::max(F * expden_k(d, eps, n), (float) num_obj);
    C.weak_phi_thr = std::max(F * alpha_k(d, eps, n), (float) num_obj * omega_k(d));
    inscy(root, C, 0);
    Run *run = new Run;
    for (auto &kv : C.result) {
        run->subspaces.push_back(kv.first);
        run->labels.push_back(kv.second);
    }
    run->trace = std::move(C.trace);
    run->calls = C.calls;
    run->trees = C.trees;
    run->queries = C.queries;
    run->nbr_total = C.nbr_total;
    return run;
}


// Set-based restatement for any SCY-tree level order (entropy_order != 0
// included): s-connection chains with their last level instead of paths.
void *oracle_sets_run_order(const float *X, int n, int d, float eps, float F, int num_obj, int min_size, float r,
                            int cells, int rectangular, int entropy_order, int threads, int trace) {
    Data D{X, n, d};
    std::vector<int> order = dimension_order(D, cells, entropy_order);
    int ecells = (int) std::ceil(1.0 / (double) eps);
    Tree etree = build_tree(D, order, ecells, eps);
    NeighbourIndex nbr(D, etree, order, false);
    if (threads > 0) omp_set_num_threads(threads);
    SetEngine E(D, nbr, eps, F, r, num_obj, min_size, cells, rectangular != 0);
    E.trace_on = trace != 0;
    E.ord = order;
    E.lvl.assign(d, 0);
    for (int t = 0; t < d; ++t) E.lvl[order[t]] = t;
    STree root;
    root.pts.resize(n);
    for (int i = 0; i < n; ++i) root.pts[i] = i;
    float cs = 1.0f / (float) cells;
    for (int x = 0; x < n; ++x)
        for (int b = 0; b < d; ++b) {
            float xb = D.row(x)[b];
            int cb = std::min((int) (xb / cs), cells - 1);
            float upper = (float) (cb + 1) * cs;
            float bound = upper - eps;
            if (xb >= bound) root.ch.push_back(Chain{b, x, d - 1});
        }
    E.rec_perm(root, 0);
    Run *run = new Run;
    for (auto &kv : E.result) {
        run->subspaces.push_back(kv.first);
        SubspaceMemo &m = E.components(kv.first);
        std::vector<int> lab(n, -1);
        for (int i = 0; i < n; ++i)
            if (m.comp[i] >= 0 && kv.second.count(m.comp[i])) lab[i] = m.comp[i];
        run->labels.push_back(lab);
    }
    run->trace = std::move(E.trace);
    run->calls = E.calls;
    run->trees = E.trees;
    run->queries = E.queries;
    run->nbr_total = 0;
    return run;
}

// Set-based restatement; same outputs as oracle_ref_run (labels are component
// root ids, i.e. equal to the reference's up to canonical relabelling).
void *oracle_sets_run(const float *X, int n, int d, float eps, float F, int num_obj, int min_size, float r,
                      int cells, int rectangular, int threads, int trace) {
    Data D{X, n, d};
    std::vector<int> order(d);
    for (int i = 0; i < d; ++i) order[i] = i;
    int ecells = (int) std::ceil(1.0 / (double) eps);
    Tree etree = build_tree(D, order, ecells, eps);
    NeighbourIndex nbr(D, etree, order, false);
    if (threads > 0) omp_set_num_threads(threads);
    SetEngine E(D, nbr, eps, F, r, num_obj, min_size, cells, rectangular != 0);
    E.trace_on = trace != 0;
    STree root;
    root.pts.resize(n);
    for (int i = 0; i < n; ++i) root.pts[i] = i;
    float cs = 1.0f / (float) cells;
    for (int x = 0; x < n; ++x)
        for (int b = 0; b < d; ++b) {
            float xb = D.row(x)[b];
            int cb = std::min((int) (xb / cs), cells - 1);
            float upper = (float) (cb + 1) * cs;
            float bound = upper - eps;
            if (xb >= bound) root.sp.push_back(SPath{b, x});
        }
    E.rec(root, 0);
    Run *run = new Run;
    for (auto &kv : E.result) {
        run->subspaces.push_back(kv.first);
        SubspaceMemo &m = E.components(kv.first);
        std::vector<int> lab(n, -1);
        for (int i = 0; i < n; ++i)
            if (m.comp[i] >= 0 && kv.second.count(m.comp[i])) lab[i] = m.comp[i];
        run->labels.push_back(lab);
    }
    run->trace = std::move(E.trace);
    run->calls = E.calls;
    run->trees = E.trees;
    run->queries = E.queries;
    run->nbr_total = 0;
    return run;
}

// CPU-baseline sample: runs the oracle for at most max_seconds of wall time
// and reports the restricted trees evaluated and the seconds taken.
void oracle_ref_rate(const float *X, int n, int d, float eps, float F, int num_obj, int min_size, float r, int cells,
                     int rectangular, double max_seconds, long long *trees, double *seconds, int *finished) {
    auto t0 = std::chrono::steady_clock::now();
    Data D{X, n, d};
    std::vector<int> order(d);
    for (int i = 0; i < d; ++i) order[i] = i;
    Tree root = build_tree(D, order, cells, eps);
    int ecells = (int) std::ceil(1.0 / (double) eps);
    Tree etree = build_tree(D, order, ecells, eps);
    NeighbourIndex nbr(D, etree, order, false);
    Ctx C{D, nbr, eps, F, r, num_obj, min_size, rectangular != 0};
    C.weak_rect_thr = std::max(F * expden_k(d, eps, n), (float) num_obj);
    C.weak_phi_thr = std::max(F * alpha_k(d, eps, n), (float) num_obj * omega_k(d));
    C.timed = true;
    C.deadline = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                          std::chrono::duration<double>(max_seconds));
    *finished = 1;
    try {
        inscy(root, C, 0);
    } catch (const Deadline &) {
        *finished = 0;
    }
    *trees = C.trees;
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int oracle_ref_count(void *h) { return (int) static_cast<Run *>(h)->subspaces.size(); }

// dims_out must hold d ints; labels_out n ints.  Returns the subspace size.
int oracle_ref_get(void *h, int i, int *dims_out, int *labels_out) {
    Run *run = static_cast<Run *>(h);
    const auto &S = run->subspaces[i];
    for (size_t k = 0; k < S.size(); ++k) dims_out[k] = S[k];
    std::memcpy(labels_out, run->labels[i].data(), run->labels[i].size() * sizeof(int));
    return (int) S.size();
}

void oracle_ref_stats(void *h, long long *out4) {
    Run *run = static_cast<Run *>(h);
    out4[0] = run->calls;
    out4[1] = run->trees;
    out4[2] = run->queries;
    out4[3] = run->nbr_total;
}

int oracle_ref_trace_len(void *h) { return (int) static_cast<Run *>(h)->trace.size(); }

// 7 int64 per row: mask, lo, hi, pre, post, pass, redundant
void oracle_ref_trace(void *h, long long *out) {
    Run *run = static_cast<Run *>(h);
    for (size_t i = 0; i < run->trace.size(); ++i) {
        const TraceRow &t = run->trace[i];
        long long *o = out + 7 * i;
        o[0] = (long long) t.mask;
        o[1] = t.lo;
        o[2] = t.hi;
        o[3] = t.pre;
        o[4] = t.post;
        o[5] = t.pass;
        o[6] = t.redundant;
    }
}

void oracle_ref_free(void *h) { delete static_cast<Run *>(h); }

// join / pruneRedundancy on hand-built inputs.  The result map is given as
// nres entries: dims (ascending, lengths in res_k), labels [nres][n].
// oracle_join_test writes the updated map (same entry order as the map's
// vec_cmp order, at most nres+1 entries) and returns the new entry count.
static Result make_result(int n, int nres, const int *res_k, const int *res_dims, const int *res_labels) {
    Result r;
    int at = 0;
    for (int e = 0; e < nres; ++e) {
        std::vector<int> S(res_dims + at, res_dims + at + res_k[e]);
        at += res_k[e];
        r.emplace(S, std::vector<int>(res_labels + (size_t) e * n, res_labels + (size_t) (e + 1) * n));
    }
    return r;
}

int oracle_join_test(int n, int nres, const int *res_k, const int *res_dims, const int *res_labels, const int *cl_in,
                     int k, const int *dims, int min_size, float r, int *out_k, int *out_dims, int *out_labels) {
    Result res = make_result(n, nres, res_k, res_dims, res_labels);
    std::vector<int> cl(cl_in, cl_in + n);
    join(res, cl, std::vector<int>(dims, dims + k), min_size, r);
    int e = 0, at = 0;
    for (const auto &kv : res) {
        out_k[e] = (int) kv.first.size();
        for (int x : kv.first) out_dims[at++] = x;
        std::memcpy(out_labels + (size_t) e * n, kv.second.data(), n * sizeof(int));
        ++e;
    }
    return e;
}

// pruneRedundancy of a region holding `pts` (its point count npoints) in subspace dims.
int oracle_redundancy_test(int n, int nres, const int *res_k, const int *res_dims, const int *res_labels,
                           const int *pts, int npts, int k, const int *dims, float r) {
    Result res = make_result(n, nres, res_k, res_dims, res_labels);
    Tree t;
    t.rdims.assign(dims, dims + k);
    t.add(-1, npts, true);
    t.nodes[0].pts.assign(pts, pts + npts);
    t.npoints = npts;
    // prune_redundancy reads only the region's points and count, C.result and C.r
    Data D{nullptr, n, 0};
    Tree et;
    et.cells = 1;
    et.cell_size = 1.0f;
    NeighbourIndex idx(D, et, {}, true);
    Ctx C{D, idx, 1.0f, 1.0f, r, 1, 1, true};
    C.result = res;
    return prune_redundancy(t, C) ? 1 : 0;
}

// Pieces exposed for unit checks against closed forms and brute force.
float oracle_unit_ball(int k) { return unit_ball(k); }
float oracle_alpha(int k, float eps, int n) { return alpha_k(k, eps, n); }
float oracle_expden(int k, float eps, int n) { return expden_k(k, eps, n); }
float oracle_omega(int k) { return omega_k(k); }

// Ordered ε-neighbour list of point p in subspace S (ascending dims).
int oracle_neighbours(const float *X, int n, int d, float eps, const int *S, int k, int p, int literal, int *out) {
    Data D{X, n, d};
    std::vector<int> order(d);
    for (int i = 0; i < d; ++i) order[i] = i;
    int ecells = (int) std::ceil(1.0 / (double) eps);
    Tree etree = build_tree(D, order, ecells, eps);
    NeighbourIndex nbr(D, etree, order, literal != 0);
    std::vector<int> SS(S, S + k);
    std::vector<int> nb = nbr.neighbours(p, SS, eps);
    for (size_t i = 0; i < nb.size(); ++i) out[i] = nb[i];
    return (int) nb.size();
}

// φ_S(p) (ClusteringCpu.cu:116-124) of the points ps[0..np) over their
// ε-neighbours in the neighbourhood tree's leaf order (literal: by walking the
// tree; otherwise the index), as the density test of the clustering uses it.
void oracle_phi_many(const float *X, int n, int d, float eps, const int *S, int k, const int *ps, int np,
                     int literal, float *out) {
    Data D{X, n, d};
    std::vector<int> order(d);
    for (int i = 0; i < d; ++i) order[i] = i;
    int ecells = (int) std::ceil(1.0 / (double) eps);
    Tree etree = build_tree(D, order, ecells, eps);
    NeighbourIndex nbr(D, etree, order, literal != 0);
    std::vector<int> SS(S, S + k);
    for (int i = 0; i < np; ++i) out[i] = phi(D, ps[i], nbr.neighbours(ps[i], SS, eps), eps, SS);
}

// Dimension order used by the SCY-tree levels (identity unless entropy_order != 0).
void oracle_dimension_order(const float *X, int n, int d, int cells, int entropy_order, int *out) {
    Data D{X, n, d};
    std::vector<int> o = dimension_order(D, cells, entropy_order);
    for (int i = 0; i < d; ++i) out[i] = o[i];
}

}  // extern "C"
