"""CPU models of the GPU photon search's bookkeeping (csrc/photon.hip,
k_photon_knn), checked against the oracle's restatement of kdtree.h.

The kernel keeps the reference's result order with two devices the oracle
does not have:
  * its heap moves one-byte slot numbers beside the distances, and a new
    photon takes the slot of the entry dropped: the root, or the spare slot k
    when the photon is itself the farthest (the first one after make_heap,
    before the radius has shrunk to the root's distance);
  * its walk stacks only far children (with their plane distances, compared
    with the radius when popped) and passes over subtrees whose 32-bit shape
    signature lacks the query's shape.
The models below restate those steps in Python and must give the oracle's
(heap_sequence, knn over the same tree) ids, distances and order exactly.
A model with the root's slot always taken fails the first test."""
import numpy as np
import pytest

import oracle


# libstdc++'s heap steps, as the kernel's heap_push_up / heap_adjust /
# heap_make / heap_push_pop run them, on parallel lists D (distances) and V
def _push_up(D, V, hole, top, d, v):
    while hole > top:
        parent = (hole - 1) // 2
        if not D[parent] < d:
            break
        D[hole], V[hole] = D[parent], V[parent]
        hole = parent
    D[hole], V[hole] = d, v


def _adjust(D, V, hole, n, d, v):
    top = child = hole
    while n > 0 and child < (n - 1) // 2:
        child = 2 * (child + 1)
        if D[child] < D[child - 1]:
            child -= 1
        D[hole], V[hole] = D[child], V[child]
        hole = child
    if n % 2 == 0 and n >= 2 and child == (n - 2) // 2:
        child = 2 * (child + 1)
        D[hole], V[hole] = D[child - 1], V[child - 1]
        hole = child - 1
    _push_up(D, V, hole, top, d, v)


def _make(D, V, n):
    for parent in range((n - 2) // 2, -1, -1):
        _adjust(D, V, parent, n, D[parent], V[parent])


def _push_pop(D, V, n, d, v):
    xd, xv = d, v
    if n > 0:
        parent = (n - 1) // 2
        if D[parent] < d:
            xd, xv = D[parent], V[parent]
            _push_up(D, V, parent, 0, d, v)
    _adjust(D, V, 0, n, xd, xv)


class SlotHeap:
    """The kernel's kHeapLdsSlot bookkeeping for one query."""

    def __init__(self, k, root_slot_always=False):
        self.k, self.D, self.S, self.pool, self.cnt = k, [0.0] * k, [0] * k, [None] * k, 0
        self.root_slot_always = root_slot_always

    def take(self, d, ident):
        k = self.k
        if self.cnt < k:
            self.D[self.cnt], self.S[self.cnt], self.pool[self.cnt] = d, self.cnt, ident
            self.cnt += 1
            if self.cnt == k:
                _make(self.D, self.S, k)
        else:
            slot = self.S[0] if (self.root_slot_always or not self.D[0] < d) else k
            if slot < k:
                self.pool[slot] = ident
            _push_pop(self.D, self.S, k, d, slot)

    def result(self):
        return [self.pool[self.S[i]] for i in range(self.cnt)], self.D[:self.cnt]


def _stream(rng, n, ties):
    d = rng.random(n).astype(np.float32)
    if ties:
        d = (np.round(d * ties) / ties).astype(np.float32)
    return d


@pytest.mark.parametrize("k", [1, 2, 3, 8, 20, 100])
def test_slot_heap_equals_the_reference_heap(k):
    rng = np.random.default_rng(k)
    for trial in range(60):
        d = _stream(rng, int(rng.integers(0, 4 * k + 40)), [None, 4, 32][trial % 3])
        ids, dist = oracle.heap_sequence(d, k)
        h = SlotHeap(k)
        for i, x in enumerate(d):
            h.take(float(x), i)
        got_ids, got_d = h.result()
        assert got_ids == ids.tolist()
        assert np.array_equal(np.float32(got_d), dist)


def test_taking_the_root_slot_always_is_caught():
    """The plausible mistake: the newcomer always takes the root's slot."""
    rng = np.random.default_rng(7)
    wrong = 0
    for _ in range(50):
        d = _stream(rng, 60, None)
        ids, _ = oracle.heap_sequence(d, 8)
        h = SlotHeap(8, root_slot_always=True)
        for i, x in enumerate(d):
            h.take(float(x), i)
        wrong += h.result()[0] != ids.tolist()
    assert wrong > 0


def _norm(a, b, c):
    s = np.float32(0.0)
    s = np.float32(s + np.float32(a * a))
    s = np.float32(s + np.float32(b * b))
    s = np.float32(s + np.float32(c * c))
    return np.float32(np.sqrt(s))


def _signatures(shape1):
    n = len(shape1)
    sig = np.zeros(n, np.int64)

    def fill(l, r):
        if r <= l:
            return 0
        med = (l + r) // 2
        s = (1 << (int(shape1[med]) & 31) if shape1[med] else 0) | fill(l, med) | fill(med + 1, r)
        sig[med] = s
        return s

    fill(0, n)
    return sig


def _walk(nodes, axis, shape1, sig, order, p, want, k, radius):
    """k_photon_knn's loop: near child next, far children stacked."""
    h = SlotHeap(k)
    maxd = np.float32(radius)
    stack = []
    l, r = 0, len(nodes)
    while True:
        if r <= l:
            while stack and not stack[-1][2] < maxd:
                stack.pop()
            if not stack:
                break
            l, r, _ = stack.pop()
        med = l + (r - l) // 2
        if want and not (sig[med] >> (want & 31)) & 1:
            r = l
            continue
        q = nodes[med]
        d = _norm(p[0] - q[0], p[1] - q[1], p[2] - q[2])
        take = d < maxd and (want == 0 or shape1[med] == want)
        if r - l > 1:
            ax = axis[med]
            split = q[ax]
            diff = np.float32(p[ax] - split)
            pl = [np.float32(0.0)] * 3
            pl[ax] = diff
            pd = _norm(*pl)
            if p[ax] < split:
                fl, fr, r = med + 1, r, med
            else:
                fl, fr, l = l, med, med + 1
            if fr > fl:
                stack.append((fl, fr, pd))
        else:
            r = l
        if take:
            full = h.cnt == k
            h.take(d, int(order[med]))
            if full:
                maxd = np.float32(h.D[0])
    return h.result()


@pytest.mark.parametrize("coplanar", [False, True])
def test_far_child_walk_with_signatures_equals_the_reference_search(coplanar):
    rng = np.random.default_rng(11 + coplanar)
    x = rng.random((2500, 3)).astype(np.float32)
    if coplanar:
        x[:1200, 2] = np.float32(0.5)  # a floor of photons: splits at equal coordinates
    # shapes are spatial cells, 40 of them so that signature bits are shared
    shape = (np.floor(x[:, 0] * 8) + 8 * np.floor(x[:, 1] * 5)).astype(np.int32)
    shape[rng.random(len(x)) < 0.03] = -1
    order, axis, _ = oracle.kd_build(x)
    nodes = x[order]
    shape1 = (shape[order] + 1).astype(np.int64)
    sig = _signatures(shape1)
    q = rng.random((160, 3)).astype(np.float32)
    q[::4, 2] = np.float32(0.5)
    qs = (np.floor(q[:, 0] * 8) + 8 * np.floor(q[:, 1] * 5)).astype(np.int32)
    qs[::5] = -1
    qs[::13] = 45  # a shape with no photon
    for k, radius in ((1, 0.2), (16, 0.15), (30, 100.0)):
        cnt, idx, dist = oracle.knn(x, q, k, radius, shape, qs, tree=(order, axis))
        for i in range(len(q)):
            ids, ds = _walk(nodes, axis, shape1, sig, order, q[i], int(qs[i]) + 1, k, radius)
            assert len(ids) == cnt[i]
            assert ids == idx[i, :cnt[i]].tolist(), (k, radius, i)
            assert np.array_equal(np.float32(ds), dist[i, :cnt[i]])
