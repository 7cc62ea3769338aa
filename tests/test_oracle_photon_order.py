"""The result order of the photon search: kdtree.h's nearest_neighbors_impl
(cuda_libraries/kdtree.h:111-152) returns its vector in insertion order
until k photons were accepted, then in the layout libstdc++'s
std::make_heap, std::push_heap and std::pop_heap leave (max-heap by
distance).  The oracle restates those steps; here they are checked against
the standard library itself: a small C++ program drives std::make_heap /
push_heap / pop_heap with the same bookkeeping over the same candidate
sequences, and both must print the same vector -- including sequences full
of equal distances, where a sift-down restatement would keep a different
photon."""
import os
import shutil
import subprocess

import numpy as np
import pytest

DRIVER = r"""
#include <algorithm>
#include <cstdio>
#include <vector>
// reads k, n, then n distances; accepts candidates 0..n-1 in order with the
// bookkeeping of a k-nearest search and prints the final vector's ids
int main() {
    unsigned k, n;
    if (std::scanf("%u %u", &k, &n) != 2) return 1;
    std::vector<float> dist(n);
    for (unsigned i = 0; i < n; ++i) std::scanf("%a", &dist[i]);
    auto cmp = [&](unsigned a, unsigned b) { return dist[a] < dist[b]; };
    std::vector<unsigned> v;
    for (unsigned i = 0; i < n; ++i) {
        v.push_back(i);
        if (v.size() == k) std::make_heap(v.begin(), v.end(), cmp);
        else if (v.size() > k) {
            std::push_heap(v.begin(), v.end(), cmp);
            std::pop_heap(v.begin(), v.end(), cmp);
            v.pop_back();
        }
    }
    for (unsigned x : v) std::printf("%u\n", x);
    return 0;
}
"""


@pytest.fixture(scope="module")
def std_heap(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("heap")
    src = d / "heap.cpp"
    src.write_text(DRIVER)
    exe = d / "heap"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe), str(src)], check=True)

    def run(dist, k):
        text = f"{k} {len(dist)}\n" + "\n".join(float(x).hex() for x in np.float32(dist)) + "\n"
        out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout
        return np.array([int(x) for x in out.split()], np.int64)
    return run


def test_advisor_example(orc):
    """Heap [10, 4, 9, 3] (already a max-heap, so make_heap keeps it) and a
    new distance 5: push_heap + pop_heap + pop_back leave [9, 5, 4, 3]; a
    replace-the-root sift-down would leave [9, 4, 5, 3]."""
    ids, d = orc.heap_sequence([10, 4, 9, 3, 5], 4)
    assert list(d) == [9, 5, 4, 3]
    assert list(ids) == [2, 4, 1, 3]


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 8, 17, 100])
def test_heap_bookkeeping_equals_libstdcxx(orc, std_heap, k):
    rng = np.random.default_rng(k)
    for trial in range(40):
        n = int(rng.integers(0, 3 * k + 10))
        if trial % 2:
            dist = rng.integers(0, 6, n).astype(np.float32)  # many ties
        else:
            dist = rng.random(n).astype(np.float32)
        ids, _ = orc.heap_sequence(dist, k)
        want = std_heap(dist, k)
        assert np.array_equal(ids, want), (k, trial, dist, ids, want)


def test_tree_search_order_is_the_heap_order_of_its_candidates(orc):
    """The oracle's kd-tree search returns the k nearest in the heap layout:
    its vector is a max-heap by distance once k photons were found, and the
    same set as brute force on unambiguous queries."""
    rng = np.random.default_rng(3)
    x = rng.random((4000, 3)).astype(np.float32)
    q = rng.random((200, 3)).astype(np.float32)
    tree = orc.kd_build(x)[:2]
    k = 16
    cnt, idx, dist = orc.knn(x, q, k, np.inf, tree=tree)
    bc, bi, bd = orc.knn(x, q, k, np.inf)
    for i in range(q.shape[0]):
        assert cnt[i] == k
        d = dist[i]
        for j in range(1, k):
            assert d[(j - 1) // 2] >= d[j]  # max-heap property
        assert sorted(idx[i]) == sorted(bi[i])
