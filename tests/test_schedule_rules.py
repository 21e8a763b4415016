"""Host checks of two ordering rules the device kernels rely on (no device).

1. The large-workspace EPA builds the horizon in parallel: with P
   occurrences of the directed edge (u, w) and N of (w, u) among the visible
   faces' edges, the survivors of the reference's find + erase on a vector
   (src/epa.hpp:95-112) are the last P - N occurrences of (u, w) when P > N,
   in order.  Checked here against the sequential list on random sequences
   with heavy reverse-edge mixing, the case of degenerate polytopes.
2. The ordered windows (k_solve_large, k_adj_island) run a sequence of visits
   on 64 lanes: a lane runs once the latest earlier lane of its window
   touching each of its bodies has run.  Checked here: no body is touched
   twice in one pass, every body's visits keep sequence order, and every
   window ends.
"""
import random

NONE = -1


def _horizon_sequential(edges):
    """The reference's list: erase the first reverse entry, else append."""
    hor = []
    for u, w in edges:
        for e, (a, b) in enumerate(hor):
            if a == w and b == u:
                del hor[e]
                break
        else:
            hor.append((u, w))
    return hor


def _horizon_by_counts(edges):
    """The rule the parallel horizon implements: keep an occurrence when its
    edge occurs P > N times against N reverse occurrences and its rank among
    the edge's occurrences is >= N."""
    count, seen, out = {}, {}, []
    for e in edges:
        count[e] = count.get(e, 0) + 1
    for u, w in edges:
        p, n = count[(u, w)], count.get((w, u), 0)
        rank = seen.get((u, w), 0)
        seen[(u, w)] = rank + 1
        if p > n and rank >= n:
            out.append((u, w))
    return out


def test_parallel_horizon_rule_equals_the_sequential_list():
    rng = random.Random(1)
    for _ in range(5000):
        nv = rng.randint(2, 6)
        edges = []
        for _ in range(rng.randint(0, 40)):
            u, w = rng.randrange(nv), rng.randrange(nv)
            if u != w:
                edges.append((u, w))
        assert _horizon_by_counts(edges) == _horizon_sequential(edges), edges


def test_parallel_horizon_rule_on_a_closed_polytope():
    """All faces of a tetrahedron visible: every edge cancels, no horizon."""
    faces = [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 3, 2)]
    edges = [e for a, b, c in faces for e in ((a, b), (b, c), (c, a))]
    assert _horizon_sequential(edges) == [] == _horizon_by_counts(edges)
    # one face visible: its three edges, in order
    assert _horizon_by_counts(edges[:3]) == [(0, 1), (1, 2), (2, 0)] == _horizon_sequential(edges[:3])


def _windows(seq, width=64):
    """Passes of the ordered-window schedule; each pass is a list of visits."""
    passes = []
    for base in range(0, len(seq), width):
        win = seq[base:base + width]
        deps = []
        for lane, (bi, bj) in enumerate(win):
            di = dj = -1
            for l in range(lane):
                oi, oj = win[l]
                if bi != NONE and bi in (oi, oj):
                    di = l
                if bj != NONE and bj in (oi, oj):
                    dj = l
            deps.append((di, dj))
        done = [False] * len(win)
        for _ in range(width + 1):
            if all(done):
                break
            ready = [l for l in range(len(win)) if not done[l]
                     and (deps[l][0] < 0 or done[deps[l][0]]) and (deps[l][1] < 0 or done[deps[l][1]])]
            assert ready, "a pass ran nothing"
            passes.append([base + l for l in ready])
            for l in ready:
                done[l] = True
        assert all(done), "a window did not end"
    return passes


def test_ordered_windows_keep_each_body_in_sequence_order():
    rng = random.Random(3)
    for _ in range(1500):
        nb = rng.randint(1, 30)
        seq = []
        for _ in range(rng.randint(0, 300)):
            a = rng.choice([NONE] + list(range(nb)))
            b = rng.choice([NONE] + list(range(nb)))
            if a == b and a != NONE:
                b = NONE
            seq.append((a, b))
        when = {}
        for r, visits in enumerate(_windows(seq)):
            touched = set()
            for v in visits:
                when[v] = r
                for x in seq[v]:
                    if x != NONE:
                        assert x not in touched, "a body touched twice in one pass"
                        touched.add(x)
        assert sorted(when) == list(range(len(seq)))
        last = {}
        for v, (a, b) in enumerate(seq):
            for x in (a, b):
                if x == NONE:
                    continue
                if x in last:
                    assert when[last[x]] < when[v], "a body's visits out of order"
                last[x] = v
