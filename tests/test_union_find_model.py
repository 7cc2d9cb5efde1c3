"""The union-find of the component kernels (inscy_device.h uf_find/uf_union)
under stale reads, restated in Python.

The kernels read parents with plain loads, so a find may see any value a
parent word held before (an L2 copy another XCD has since changed); only the
linking compare-and-swap sees the current value.  This model is more
adversarial than the hardware: every read returns a randomly chosen past
value of the word.  Over many random interleavings of concurrent unions it
checks that every find and union ends within a bound and that the final
forest is exactly the connected components of the edges (brute force).
A planted mistake (giving up when the swap fails instead of retrying with
the value it returns) must fail it.
"""

import random

import pytest


class Words:
    """par as a history per word: reads may return any past value."""

    def __init__(self, n, rng):
        self.hist = [[i] for i in range(n)]
        self.rng = rng

    def load(self, x):  # plain load: any value the word ever held
        return self.rng.choice(self.hist[x])

    def store(self, x, v):  # plain store
        self.hist[x].append(v)

    def cas(self, x, expect, v):  # device-scope compare-and-swap: current value
        cur = self.hist[x][-1]
        if cur == expect:
            self.hist[x].append(v)
        return cur

    def current(self, x):
        return self.hist[x][-1]


def uf_find(W, x):
    """inscy_device.h uf_find, one memory operation per yield."""
    p = W.load(x)
    yield
    while p != x:
        g = W.load(p)
        yield
        if g != p:
            W.store(x, g)  # path halving
            yield
        x, p = p, g
    return x


def uf_union(W, a, b, planted=False):
    """inscy_device.h uf_union (planted: no retry after a failed swap)."""
    while True:
        a = yield from uf_find(W, a)
        b = yield from uf_find(W, b)
        if a == b:
            return
        if a > b:
            a, b = b, a
        old = W.cas(b, b, a)
        yield
        if old == b or planted:
            return
        b = old


def components(n, edges):
    par = list(range(n))

    def f(x):
        while par[x] != x:
            x = par[x]
        return x

    for a, b in edges:
        ra, rb = f(a), f(b)
        if ra != rb:
            par[max(ra, rb)] = min(ra, rb)
    return [f(x) for x in range(n)]


def run_schedule(seed, n=40, m=70, threads=8, planted=False, limit=200000):
    rng = random.Random(seed)
    edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(m)]
    W = Words(n, rng)
    work = [edges[t::threads] for t in range(threads)]
    gens = []
    for t in range(threads):
        def thread(es=work[t]):
            for a, b in es:
                yield from uf_union(W, a, b, planted)
        gens.append(thread())
    steps = 0
    while gens:
        g = rng.choice(gens)
        try:
            next(g)
        except StopIteration:
            gens.remove(g)
        steps += 1
        if steps > limit:
            return None, edges, W
    return steps, edges, W


def final_roots(W, n):
    out = []
    for x in range(n):
        while W.current(x) != x:
            x = W.current(x)
        out.append(x)
    return out


@pytest.mark.parametrize("seed", range(60))
def test_concurrent_unions_with_stale_reads_give_the_components(seed):
    steps, edges, W = run_schedule(seed)
    assert steps is not None, "a union did not end"
    n = len(W.hist)
    # every value a word ever held is at most the word's own id (no cycles)
    assert all(v <= x for x in range(n) for v in W.hist[x])
    assert final_roots(W, n) == components(n, edges)


def test_planted_mistake_is_caught():
    # a union whose swap lost a race must go on with the current root;
    # giving up loses the edge in some schedule
    wrong = 0
    for seed in range(60):
        steps, edges, W = run_schedule(seed, planted=True)
        n = len(W.hist)
        wrong += final_roots(W, n) != components(n, edges)
    assert wrong > 0
