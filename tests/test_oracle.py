"""Pins the oracle (the CPU restatement of the reference's CPU INSCY) against
what fixes it independently of its own code: closed forms, brute force, a
literal walk of the neighbourhood tree, a second restatement (set-based) of
the whole algorithm, planted mistakes that the cross-check must catch, and the
committed golden fixture of the first workload."""

import math
import os
import subprocess
import sys

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

from gpu_inscy_amd import datagen
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------------ thresholds
def ball_closed_form(k):
    # c(k) = PI^(k/2) / Gamma(k/2 + 1) with the reference's PI = 3.14 used both in
    # the power and in Gamma(1/2) = sqrt(PI) (ClusteringCpu.cu:25,136-150):
    # even k: PI^(k/2) / (k/2)!;  odd k: PI^((k-1)/2) 2^((k+1)/2) / k!!
    if k % 2 == 0:
        return 3.14 ** (k // 2) / math.factorial(k // 2)
    double_fact = math.prod(range(k, 0, -2))
    return 3.14 ** ((k - 1) // 2) * 2 ** ((k + 1) // 2) / double_fact


# The reference's rounding sequence (ClusteringCpu.cu:36-38,127-164), written
# from that source with explicit float32/float64 steps: gamma() returns a float
# of a double product, c() divides two floats, alpha()/expDen() round double
# products of C pow() to float, omega() rounds a double quotient.
def ref_gamma(m):
    if m == 2:
        return np.float32(1.0)
    if m == 1:
        return np.float32(math.sqrt(3.14))
    return np.float32((m / 2.0 - 1.0) * float(ref_gamma(m - 2)))


def ref_c(k):
    return np.float32(np.float32(math.pow(3.14, k / 2.0)) / ref_gamma(k + 2))


def ref_alpha(k, eps, n):
    e = float(np.float32(eps))
    r = np.float32(float(2 * n) * math.pow(e, k) * float(ref_c(k)))
    return np.float32(float(r) / (math.pow(1.0, k) * (k + 2)))


def ref_expden(k, eps, n):
    nc = np.float32(np.float32(n) * ref_c(k))
    r = np.float32(float(nc) * math.pow(float(np.float32(eps)), k))
    return np.float32(float(r) / math.pow(1.0, k))


def ref_omega(k):
    return np.float32(2.0 / (k + 2.0))


def exact_ball(k):
    """c(k) = 3.14^(k/2) / Gamma(k/2 + 1) with Gamma(1/2) = sqrt(3.14), to 50 digits."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    pi = Decimal("3.14")
    if k % 2 == 0:
        return float(pi ** (k // 2) / math.factorial(k // 2))
    # odd k: sqrt(pi) cancels; c(k) = pi^((k-1)/2) 2^((k+1)/2) / k!!
    return float(pi ** ((k - 1) // 2) * 2 ** ((k + 1) // 2) / math.prod(range(k, 0, -2)))


def ulps(a, b):
    a, b = np.float32(a), np.float32(b)
    return abs(int(a.view(np.int32)) - int(b.view(np.int32)))


THRESHOLD_CASES = [(k, eps, n) for k in range(0, 33) for eps, n in
                   [(0.01, 1500), (0.015, 25000), (0.02, 100000), (0.02, 200000), (0.003, 7), (0.25, (1 << 30) - 1)]]  # n <= 2^30 - 1: 2 n fits an int


@pytest.mark.parametrize("k", list(range(1, 33)))
def test_unit_ball_volume_bit_exact_and_within_the_closed_form(k):
    got = np.float32(O.thresholds(k, 0.01, 1000)["c"])
    # bit-exact: the reference's float/double sequence
    assert got.view(np.int32) == ref_c(k).view(np.int32)
    # and within the rounding of that sequence of the exact value: one
    # rounding per gamma step, the pow, the division
    assert ulps(got, exact_ball(k)) <= 3 + k // 2


def test_thresholds_bit_exact_for_every_k_eps_n():
    for k, eps, n in THRESHOLD_CASES:
        t = O.thresholds(k, eps, n)
        for name, want in (("c", ref_c(k)), ("alpha", ref_alpha(k, eps, n)), ("expden", ref_expden(k, eps, n)),
                           ("omega", ref_omega(k))):
            got = np.float32(t[name])
            assert got.view(np.int32) == want.view(np.int32), (name, k, eps, n, got, want)


def test_threshold_check_catches_a_wrong_float_double_order():
    # the bit-exact check is discriminating: computing c(k) wholly in double,
    # or Gamma in float arithmetic, changes the float result for some k
    def c_double(k):
        return np.float32(math.pow(3.14, k / 2.0) / float(math.gamma(k / 2.0 + 1) / math.sqrt(math.pi)
                                                           * (math.sqrt(3.14) if k % 2 else 1.0)))

    # (Gamma's float recursion itself is no such mistake: each product of a
    # half-integer and a float is exact in double, so it rounds once either way)
    def gamma_true_pi(m):
        if m == 2:
            return np.float32(1.0)
        if m == 1:
            return np.float32(math.sqrt(math.pi))
        return np.float32((m / 2.0 - 1.0) * float(gamma_true_pi(m - 2)))

    def c_true_pi_gamma(k):
        return np.float32(np.float32(math.pow(3.14, k / 2.0)) / gamma_true_pi(k + 2))

    assert any(c_double(k) != ref_c(k) for k in range(1, 33))
    assert any(c_true_pi_gamma(k) != ref_c(k) for k in range(1, 33))

    def expden_all_float(k, eps, n):
        return np.float32(np.float32(n) * ref_c(k) * np.float32(eps) ** np.float32(k))

    assert any(expden_all_float(k, e, n) != ref_expden(k, e, n) for k, e, n in THRESHOLD_CASES if k > 0)


def test_unit_ball_uses_pi_314_not_math_pi():
    got = O.thresholds(2, 0.01, 1000)["c"]
    assert got == pytest.approx(3.14, rel=1e-7)
    assert abs(got - math.pi) > 1e-3
    assert O.thresholds(1, 0.01, 1000)["c"] == 2.0  # PI cancels in one dimension
    assert O.thresholds(3, 0.01, 1000)["c"] == pytest.approx(4 * 3.14 / 3, rel=1e-7)


@pytest.mark.parametrize("k,eps,n", [(1, 0.01, 1500), (2, 0.015, 25000), (5, 0.02, 100000), (15, 0.015, 25000),
                                     (20, 0.02, 100000)])
def test_density_thresholds_within_rounding_of_the_closed_form(k, eps, n):
    # expDen = n c(k) eps^k and alpha = 2 n eps^k c(k) / (k + 2), eps taken as
    # its float32 value; a few float roundings beyond those of c(k)
    t = O.thresholds(k, eps, n)
    e32 = float(np.float32(eps))
    assert ulps(t["expden"], n * exact_ball(k) * e32 ** k) <= 6 + k // 2
    assert ulps(t["alpha"], 2 * n * e32 ** k * exact_ball(k) / (k + 2)) <= 6 + k // 2
    assert t["omega"] == np.float32(2.0 / (k + 2))


def test_engine_thresholds_equal_the_oracle_bit_for_bit():
    # the engine's host copies (no device call)
    import ctypes
    from gpu_inscy_amd import _lib
    L = _lib.lib()
    out = (ctypes.c_float * 4)()
    for k, eps, n in THRESHOLD_CASES:
        assert L.inscy_debug_thresholds(k, eps, n, out) == 0
        t = O.thresholds(k, eps, n)
        got = np.array(out[:], dtype=np.float32)
        want = np.array([t["c"], t["alpha"], t["expden"], t["omega"]], dtype=np.float32)
        assert np.array_equal(got.view(np.int32), want.view(np.int32)), (k, eps, n, got, want)


# ------------------------------------------------------------------ neighbourhood
def brute_neighbours(X, eps, S, p):
    """O(n) brute force of the reference's rule: eps-grid cells differ by <= 1 in
    every dim of S, float32 sum of squares in S order, sqrt <= eps, q != p."""
    X = np.asarray(X, dtype=np.float32)
    ec = int(math.ceil(1.0 / float(np.float32(eps))))
    ecs = np.float32(1.0) / np.float32(ec)
    cells = np.minimum((X / ecs).astype(np.int64), ec - 1)
    ok = np.ones(len(X), dtype=bool)
    for k in S:
        ok &= np.abs(cells[:, k] - cells[p, k]) <= 1
    acc = np.zeros(len(X), dtype=np.float32)
    for k in S:
        diff = (X[p, k] - X[:, k]).astype(np.float32)
        acc = (acc + (diff * diff).astype(np.float32)).astype(np.float32)
    dist = np.sqrt(acc).astype(np.float32)
    ok &= dist <= np.float32(eps)
    ok[p] = False
    return set(np.flatnonzero(ok).tolist())


@pytest.mark.parametrize("seed", range(6))
def test_neighbours_index_equals_literal_walk_and_brute_force(seed):
    rng = np.random.default_rng(seed)
    w = datagen.small(seed, n=400, d=5, n_clusters=2, eps=0.05, sigma=0.03)
    for _ in range(8):
        k = int(rng.integers(1, 6))
        S = sorted(rng.choice(5, size=k, replace=False).tolist())
        p = int(rng.integers(0, w.n))
        eps = float(rng.choice([0.01, 0.05, 0.1]))
        a = O.neighbours(w.X, eps, S, p, literal=False)
        b = O.neighbours(w.X, eps, S, p, literal=True)
        assert a.tolist() == b.tolist()  # same list, same order (tree traversal order)
        assert set(a.tolist()) == brute_neighbours(w.X, eps, S, p)


def test_grid_rule_misses_pairs_the_plain_distance_would_keep():
    # eps = 0.01 gives ceil(1/eps) = 101 cells of width < eps: a pair 0.0099.. apart
    # across two cell boundaries is within eps but not a candidate
    X = np.array([[0.02966], [0.03961], [0.5]], dtype=np.float32)
    w = np.float32(1.0) / np.float32(101)
    assert int(X[0, 0] / w) == 2 and int(X[1, 0] / w) == 4
    assert abs(np.float32(X[1, 0] - X[0, 0])) <= np.float32(0.01)
    assert O.neighbours(X, 0.01, [0], 0).tolist() == []


def test_dimension_order_entropy():
    rng = np.random.default_rng(3)
    X = rng.random((500, 4)).astype(np.float32)
    X[:, 2] = np.where(rng.random(500) < 0.9, X[:, 2] * 0.3, X[:, 2]).astype(np.float32)  # skewed, no empty cell
    assert O.dimension_order(X, 3, 0).tolist() == [0, 1, 2, 3]
    order = O.dimension_order(X, 3, 1).tolist()
    assert order[0] == 2 and sorted(order) == [0, 1, 2, 3]
    assert O.dimension_order(X, 3, -1).tolist() == order[::-1]
    # an empty cell makes the reference's entropy 0*log(0) = NaN
    # (inscy_map.cpp:94-95); the order is still a permutation
    X[:, 1] = np.float32(0.1)
    assert sorted(O.dimension_order(X, 3, 1).tolist()) == [0, 1, 2, 3]


# ------------------------------------------------------------------ two restatements agree
def fuzz_case(seed):
    rng = np.random.default_rng(seed + 99991)
    n = int(rng.integers(3, 16))
    d = int(rng.integers(2, 6))
    cells = int(rng.integers(2, 5))
    eps = float(rng.choice([0.05, 0.1, 0.2, 0.3, 0.5]))
    mu = int(rng.integers(1, 4))
    ms = int(rng.integers(1, 4))
    rect = bool(rng.integers(0, 2))
    X = (rng.integers(0, 2 * cells + 1, size=(n, d)) / (2.0 * cells)).astype(np.float32)
    X += (rng.random((n, d)) * 0.02).astype(np.float32)
    X = np.clip(X, 0, 1).astype(np.float32)
    X[0] = 0.0
    X[1] = 1.0
    return X, eps, mu, ms, cells, rect


def small_case(seed):
    rng = np.random.default_rng(seed + 500)
    n = int(rng.integers(60, 300))
    d = int(rng.integers(2, 6))
    cells = int(rng.integers(2, 6))
    eps = float(rng.choice([0.02, 0.05, 0.1, 0.15]))
    mu = int(rng.integers(1, 6))
    ms = int(rng.integers(1, max(2, n // 8)))
    rect = bool(rng.integers(0, 2))
    r = float(rng.choice([1.0, 0.5, 1.5]))
    w = datagen.small(seed, n=n, d=d, n_clusters=int(rng.integers(1, 4)), eps=eps, cells=cells, num_obj=mu,
                      min_size=ms, rectangular=rect, sigma=float(rng.choice([0.02, 0.05, 0.1])))
    return w.X, eps, mu, ms, cells, rect, r


def agree(X, eps, mu, ms, cells, rect, r=1.0):
    a = O.run_ref(X, eps, 1.0, mu, ms, r, cells, rect, 0, trace=True)
    b = O.run_sets(X, eps, 1.0, mu, ms, r, cells, rect, threads=1, trace=True)
    ok, msg = O.same_result(a, b)
    return ok and np.array_equal(a.trace, b.trace) and a.calls == b.calls, msg


def test_pointer_tree_and_set_restatements_agree_tiny():
    bad = [s for s in range(400) if not agree(*fuzz_case(s))[0]]
    assert not bad, f"seeds {bad[:10]}"


def test_pointer_tree_and_set_restatements_agree_small():
    bad = [s for s in range(60) if not agree(*small_case(s))[0]]
    assert not bad, f"seeds {bad[:10]}"


def test_pointer_tree_and_set_restatements_agree_beyond_64_dimensions():
    # d = 100: subspaces, s-connection paths and near-edge marks with
    # dimension ids beyond 64 (both restatements hold them in 128 bits)
    rect = True
    w = datagen.wide(200, 100, cells=2)
    a = O.run_ref(w.X, w.eps, 1.0, w.num_obj, w.min_size, 1.0, w.cells, rect, 0, trace=True)
    b = O.run_sets(w.X, w.eps, 1.0, w.num_obj, w.min_size, 1.0, w.cells, rect, threads=2, trace=True)
    ok, msg = O.same_result(a, b)
    assert ok, msg
    assert np.array_equal(a.trace, b.trace) and a.calls == b.calls
    assert any(len(s) > 1 and max(s) >= 64 for s in a.subspaces)


def test_set_restatement_refuses_prefix_keys_beyond_128_bits():
    X = np.full((4, 65), 0.5, dtype=np.float32)
    with pytest.raises(ValueError):
        O.run_sets(X, 0.1, 1.0, 1, 1, 1.0, 4, True)


@pytest.mark.parametrize("mutation", [1, 3])
def test_cross_check_catches_planted_mistakes(mutation):
    # the set restatement run with a planted mistake in its s-connection prune
    # (1: paths never pruned, 3: the sibling rule ignored) must disagree with the
    # pointer-tree restatement on some of the same cases
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from tests.test_oracle import agree, fuzz_case\n"
            "print(sum(1 for s in range(300) if not agree(*fuzz_case(s))[0]))\n") % ROOT
    env = dict(os.environ, ORACLE_SETS_MUTATE=str(mutation))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert int(out.stdout.strip().splitlines()[-1]) > 0


def test_simplified_path_rule_is_exact():
    # the engine's rule (a path survives iff a surviving point shares its cells
    # on every unrestricted dimension up to its branch) equals the reference's
    # three-part rule on every case
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from tests.test_oracle import agree, fuzz_case, small_case\n"
            "bad = sum(1 for s in range(300) if not agree(*fuzz_case(s))[0])\n"
            "bad += sum(1 for s in range(40) if not agree(*small_case(s))[0])\n"
            "print(bad)\n") % ROOT
    env = dict(os.environ, ORACLE_SETS_MUTATE="5")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert int(out.stdout.strip().splitlines()[-1]) == 0


# ------------------------------------------------------------------ clusters are dense components
def test_clusters_are_components_of_the_dense_graph():
    w = datagen.config(0)
    ref = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0)
    n = w.n
    for S, lab in zip(ref.subspaces, ref.labels):
        lists = [brute_neighbours(w.X, w.eps, S, p) for p in range(n)]
        thr = max(w.F * O.thresholds(len(S), w.eps, n)["expden"], float(w.num_obj))
        dense = np.array([np.float32(len(l)) >= np.float32(thr) for l in lists])
        rows, cols = [], []
        for p in range(n):
            if dense[p]:
                for q in lists[p]:
                    if dense[q]:
                        rows.append(p)
                        cols.append(q)
        g = csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
        _, comp = connected_components(g, directed=False)
        canon = O.canonical(lab)
        for c in range(canon.max() + 1):
            members = np.flatnonzero(canon == c)
            assert dense[members].all()
            comp_ids = set(comp[members].tolist())
            assert len(comp_ids) == 1
            whole = np.flatnonzero((comp == comp[members[0]]) & dense)
            assert np.array_equal(np.sort(members), whole)
            assert len(members) >= w.min_size


# ------------------------------------------------------------------ golden fixture
def test_oracle_matches_config0_golden():
    g = np.load(os.path.join(ROOT, "tests", "golden", "config0.npz"))
    w = datagen.config(0)
    ref = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0, trace=True)
    masks = [sum(1 << k for k in s) for s in ref.subspaces]
    assert masks == g["subspace_masks"].tolist()
    assert np.array_equal(np.stack([O.canonical(l) for l in ref.labels]), g["labels"])
    assert np.array_equal(ref.trace, g["trace"])


def test_literal_walk_oracle_matches_index_oracle_on_config0():
    w = datagen.config(0)
    a = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0, literal=False, trace=True)
    b = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, 0, literal=True, trace=True)
    assert O.same_result(a, b)[0] and np.array_equal(a.trace, b.trace)


# ------------------------------------------------------------------ phi density
def test_phi_hand_computed_on_dyadic_points():
    # eps = 0.5 (two eps-cells, all adjacent); every distance, ratio and square
    # below is a short dyadic fraction, so phi is exact:
    #   p = 0 in S = {0}: q at 0.25 -> r 0.5 -> 0.75; q at 0.125 -> 0.9375;
    #   q at 0.5 -> r 1 -> 0; q at 0.875 is beyond eps        => 1.6875
    #   p = 0 in S = {0, 1}: q = (0.1875, 0.25) at 0.3125 -> r 0.625 -> 0.609375;
    #   q = (0.125, 0) -> 0.9375; q = (0.25, 0.5) at 0.5590.. -> out   => 1.546875
    X = np.array([[0.0, 0.0], [0.25, 0.75], [0.125, 0.0], [0.5, 0.875], [0.875, 0.0], [0.1875, 0.25],
                  [0.25, 0.5]], dtype=np.float32)
    # S = {0}: neighbours of 0 are 1 (0.25), 2 (0.125), 3 (0.5), 5 (0.1875), 6 (0.25)
    want0 = 0.75 + 0.9375 + 0.0 + (1 - 0.375 ** 2) + 0.75
    assert O.phi_many(X, 0.5, [0], [0])[0] == np.float32(want0)
    assert O.phi_many(X, 0.5, [0, 1], [0])[0] == np.float32(0.609375 + 0.9375)
    for lit in (False, True):
        assert O.phi_many(X, 0.5, [1], [4], literal=lit)[0] == np.float32(1.0 + 1.0 + 0.75)  # 0, 2 at 0; 5 at 0.25


def ecells_of(X, eps):
    ec = int(math.ceil(1.0 / float(np.float32(eps))))
    ecs = np.float32(1.0) / np.float32(ec)
    return np.minimum((X / ecs).astype(np.int64), ec - 1)


def phi_float32_leaf_order(X, eps, S, p):
    """ClusteringCpu.cu:116-124 in numpy: terms 1 - (dist/eps)^2 with the
    distance, ratio and square in float, added to a float sum through double,
    over the neighbours in the neighbourhood tree's leaf order: ascending
    eps-cells over all dimensions (level 0 first), then insertion (id) order."""
    X = np.asarray(X, dtype=np.float32)
    nb = sorted(brute_neighbours(X, eps, S, p))
    cells = ecells_of(X, eps)
    nb.sort(key=lambda q: (tuple(cells[q].tolist()), q))
    e = np.float32(eps)
    acc = np.float32(0.0)
    for q in nb:
        s2 = np.float32(0.0)
        for k in S:
            diff = np.float32(X[p, k] - X[q, k])
            s2 = np.float32(s2 + np.float32(diff * diff))
        r = np.float32(np.float32(np.sqrt(s2)) / e)
        sq = np.float32(r * r)
        acc = np.float32(float(acc) + (1.0 - float(sq)))
    return acc, nb


@pytest.mark.parametrize("seed,S,eps", [(0, [0], 0.05), (1, [0, 2], 0.08), (2, [1, 2, 3], 0.12), (3, [3], 0.03)])
def test_phi_bit_exact_against_float32_leaf_order_emulation(seed, S, eps):
    w = datagen.small(seed, n=500, d=4, n_clusters=2, eps=eps, sigma=0.03)
    X = w.X
    ps = np.arange(0, len(X), 7, dtype=np.int32)
    got = O.phi_many(X, eps, S, ps)
    lit = O.phi_many(X, eps, S, ps[:20], literal=True)
    assert np.array_equal(lit.view(np.int32), got[:20].view(np.int32))
    order_matters = 0
    for p, g in zip(ps, got):
        want, nb = phi_float32_leaf_order(X, eps, S, int(p))
        assert np.float32(g).view(np.int32) == want.view(np.int32), (p, g, want)
        # the same terms summed in reverse give another float for some points:
        # the check pins the order, not only the terms
        rev = np.float32(0.0)
        for q in reversed(nb):
            d2 = np.float32(0.0)
            for k in S:
                diff = np.float32(X[p, k] - X[q, k])
                d2 = np.float32(d2 + np.float32(diff * diff))
            r = np.float32(np.float32(np.sqrt(d2)) / np.float32(eps))
            rev = np.float32(float(rev) + (1.0 - float(np.float32(r * r))))
        order_matters += int(rev != want)
    assert order_matters > 0 or len(S) == 0


@pytest.mark.parametrize("seed,S,eps", [(4, [0, 1], 0.1), (5, [2], 0.04), (6, [0, 1, 2, 3], 0.2)])
def test_phi_within_the_double_sum_band(seed, S, eps):
    # float64 brute force: the float phi is within a relative N 2^-22 of the
    # double sum of the same float terms in any order (the band the engine's
    # double-sum scan relies on to decide thresholds), and within N 2^-20 of
    # the phi of float64 distances (a dropped term or a wrong sign is not)
    w = datagen.small(seed, n=800, d=4, n_clusters=3, eps=eps, sigma=0.04)
    X = w.X
    ps = np.arange(0, len(X), 5, dtype=np.int32)
    got = O.phi_many(X, eps, S, ps).astype(np.float64)
    X64 = X.astype(np.float64)
    for p, g in zip(ps, got):
        nb = np.array(sorted(brute_neighbours(X, eps, S, int(p))), dtype=np.int64)
        N = len(nb)
        if N == 0:
            assert g == 0.0
            continue
        diff = (X[p, S][None, :] - X[nb][:, S]).astype(np.float32)
        s2 = np.zeros(N, dtype=np.float32)
        for t in range(len(S)):
            s2 = (s2 + (diff[:, t] * diff[:, t]).astype(np.float32)).astype(np.float32)
        r = (np.sqrt(s2).astype(np.float32) / np.float32(eps)).astype(np.float32)
        terms = 1.0 - (r * r).astype(np.float32).astype(np.float64)
        dsum = float(np.sum(terms[np.random.default_rng(int(p)).permutation(N)]))
        assert abs(g - dsum) <= N * 2.0 ** -22 * dsum + 1e-12
        d64 = np.sqrt(((X64[p, S][None, :] - X64[nb][:, S]) ** 2).sum(axis=1)) / float(np.float32(eps))
        phi64 = float(np.sum(1.0 - d64 ** 2))
        assert abs(g - phi64) <= N * 2.0 ** -20 * max(phi64, 1.0)


def test_non_rectangular_mode_oracles_agree_on_config0_shape():
    w = datagen.config(0)
    a = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, False, 0, trace=True)
    b = O.run_sets(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, False, threads=2, trace=True)
    assert O.same_result(a, b)[0]


# ------------------------------------------------------------------ join / pruneRedundancy, hand-built
def test_join_nothing_labelled_leaves_result():
    res = [([0, 1], [0, 0, -1, -1])]
    assert O.join_test(res, [-1, -2, -1, -1], [0], 1) == res


def test_join_drops_small_clusters_then_inserts():
    out = O.join_test([], [0, 0, 0, 1, -1], [2], 2)
    assert out == [([2], [0, 0, 0, -1, -1])]


def test_join_redundancy_against_superspace_depends_on_r():
    res = [([0, 1], [5, 5, 5, 5, -1])]  # superspace cluster of 4 holding the new cluster's points
    cl = [0, 0, 0, -1, -1]               # cluster of 3 in S = {0}
    assert O.join_test(res, cl, [0], 1, r=1.0) == res            # 3 * 1.0 < 4: removed, nothing inserted
    assert O.join_test(res, cl, [0], 1, r=1.5) == [([0], cl), ([0, 1], [5, 5, 5, 5, -1])]  # 4.5 >= 4: kept
    # {1, 2} is not a superspace of {0}: no removal
    res2 = [([1, 2], [5, 5, 5, 5, -1])]
    assert O.join_test(res2, cl, [0], 1) == [([0], cl), ([1, 2], [5, 5, 5, 5, -1])]


def test_join_merges_into_existing_entry_with_offset_labels():
    res = [([1], [0, 0, -1, -1, -1])]    # max label m = 0
    cl = [-1, -2, -1, 3, 3]
    out = O.join_test(res, cl, [1], 1)
    # cl >= 0 -> m + 1 + cl = 4; cl == -2 copied; others keep the old label
    assert out == [([1], [0, -2, -1, 4, 4])]


def test_join_result_order_is_ascending_bitmask():
    res = [([0, 1], [0, -1, -1]), ([2], [0, -1, -1]), ([1], [0, -1, -1])]
    out = O.join_test(res, [0, 0, -1], [0, 2], 1, r=10.0)
    assert [s for s, _ in out] == [[1], [0, 1], [2], [0, 2]]
    # beyond 64 dimensions the largest differing dimension decides (vec_cmp)
    res = [([70], [0, -1, -1]), ([0, 69], [0, -1, -1]), ([5, 100], [0, -1, -1])]
    out = O.join_test(res, [0, 0, -1], [64], 1, r=10.0)
    assert [s for s, _ in out] == [[64], [0, 69], [70], [5, 100]]


def test_redundancy_hand_built():
    res = [([0, 1], [7, 7, 7, 9, 9, -1])]   # cluster 7 of 3, cluster 9 of 2
    assert not O.redundancy_test(res, 6, [3, 4], [0], r=1.0)   # 2 * 1.0 > 2 is false
    assert O.redundancy_test(res, 6, [3, 4], [0], r=1.5)       # 3 > 2
    assert not O.redundancy_test(res, 6, [0, 3], [0], r=1.0)   # min over touched clusters = 2
    assert O.redundancy_test(res, 6, [5], [0], r=1.0)           # touches no cluster: keep
    assert O.redundancy_test(res, 6, [3, 4], [2], r=1.0)       # {0, 1} is not a superspace of {2}
    res2 = res + [([0, 2], [1, 1, 1, 1, 1, -1])]              # second superspace, min 5
    assert not O.redundancy_test(res2, 6, [0, 1, 2, 3], [0], r=1.0)   # max of mins = 5 >= 4
    assert O.redundancy_test(res2, 6, [0, 1, 2, 3, 4, 5], [0], r=1.0)  # 6 > 5


# ------------------------------------------------------------------ permuted level order
@pytest.mark.parametrize("seed", range(40))
def test_set_restatement_of_a_permuted_level_order_matches_the_pointer_tree(seed):
    # the chain model (truncated s-connection chains, empty nodes above the
    # restricted level, mergeNodes' positive-count rule) against the literal
    # pointer-tree restatement, both density modes, ascending and descending
    # entropy order; per-tree traces included
    rng = np.random.default_rng(10_000 + seed)
    d = int(rng.integers(2, 8)); n = int(rng.integers(80, 400)); cells = int(rng.integers(2, 7))
    eps = float(rng.choice([0.02, 0.04, 0.06, 0.1, 0.15]))
    rect = bool(rng.integers(0, 2))
    w = datagen.small(10_000 + seed, n=n, d=d, n_clusters=int(rng.integers(1, 5)), eps=eps, cells=cells,
                      num_obj=int(rng.integers(2, 6)), min_size=int(rng.integers(3, 20)),
                      sigma=float(rng.choice([0.02, 0.04, 0.07, 0.12])), sub_dims=(1, min(d, 4)))
    for order in (1, -1, 0):
        a = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, rect, order, trace=True)
        b = O.run_sets_order(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, rect, order, threads=1, trace=True)
        assert O.same_result(a, b)[0], (order, O.same_result(a, b)[1])
        assert np.array_equal(a.trace, b.trace)
        assert (a.calls, a.trees) == (b.calls, b.trees)


def permuted_case(seed):
    rng = np.random.default_rng(10_000 + seed)
    d = int(rng.integers(2, 8)); n = int(rng.integers(80, 400)); cells = int(rng.integers(2, 7))
    eps = float(rng.choice([0.02, 0.04, 0.06, 0.1, 0.15]))
    w = datagen.small(10_000 + seed, n=n, d=d, n_clusters=int(rng.integers(1, 5)), eps=eps, cells=cells,
                      num_obj=int(rng.integers(2, 6)), min_size=int(rng.integers(3, 20)),
                      sigma=float(rng.choice([0.02, 0.04, 0.07, 0.12])), sub_dims=(1, min(d, 4)))
    return w


def permuted_agree(seed, order):
    w = permuted_case(seed)
    a = O.run_ref(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, order, trace=True)
    b = O.run_sets_order(w.X, w.eps, w.F, w.num_obj, w.min_size, w.r, w.cells, True, order, threads=1, trace=True)
    return O.same_result(a, b)[0] and np.array_equal(a.trace, b.trace)


@pytest.mark.parametrize("mutation", [6, 7, 8])
def test_permuted_cross_check_catches_planted_mistakes(mutation):
    # the set restatement of a permuted level order run with a planted mistake
    # (6: mergeNodes' positive-count rule ignored, 7: the sibling rule ignored,
    # 8: the empty nodes kept above the restricted level not counted as
    # existing siblings) must disagree with the pointer tree on some cases
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from tests.test_oracle import permuted_agree\n"
            "print(sum(1 for s in range(120) for o in (1, -1) if not permuted_agree(s, o)))\n") % ROOT
    env = dict(os.environ, ORACLE_SETS_MUTATE=str(mutation))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    assert int(out.stdout.strip().splitlines()[-1]) > 0
