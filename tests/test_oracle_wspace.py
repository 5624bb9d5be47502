"""The oracle's workspace Sort / Merge (utils/wspace.h:62-217), pinned
against independent references: numpy's lexicographic sort and a dictionary
sum over coordinates."""
import numpy as np

import oracle
from helpers import random_records, sorted_coo


def keys_of(w):
    return w["crd"][:, 0].astype(np.int64) * (1 << 32) + w["crd"][:, 1].astype(np.int64)


def test_sort_ascending_matches_lexsort():
    rng = np.random.default_rng(0)
    w = random_records(rng, 5000, 300, 400)
    s, n = oracle.wspace_sort(w, rev=False)
    assert n == w.size
    order = np.lexsort((w["crd"][:, 1], w["crd"][:, 0]))
    np.testing.assert_array_equal(s, w[order])


def test_sort_descending_counts_valid_records():
    rng = np.random.default_rng(1)
    w = random_records(rng, 4000, 300, 400, empty_frac=0.3)
    s, n = oracle.wspace_sort(w, rev=True)
    valid = w[w["crd"][:, 0] != -1]
    assert n == valid.size
    order = np.lexsort((valid["crd"][:, 1], valid["crd"][:, 0]))[::-1]
    np.testing.assert_array_equal(s[:n], valid[order])
    assert np.all(s["crd"][n:, 0] == -1)


def test_sort_descending_all_empty():
    w = np.zeros(7, dtype=oracle.WSPACE_DTYPE)
    w["crd"][:, 0] = -1
    s, n = oracle.wspace_sort(w, rev=True)
    assert n == 7 and np.all(s["crd"][:, 0] == -1)


def dict_merge(c0, c1, v, w):
    acc = {}
    for a, b, x in zip(c0.tolist(), c1.tolist(), v.tolist()):
        acc[(a, b)] = acc.get((a, b), 0.0) + x
    for (a, b), x in zip(w["crd"].tolist(), w["val"].tolist()):
        acc[(a, b)] = acc.get((a, b), 0.0) + x
    ks = sorted(acc)
    return (np.array([k[0] for k in ks], np.int32), np.array([k[1] for k in ks], np.int32),
            np.array([acc[k] for k in ks]))


def test_merge_both_directions_match_dictionary_sum():
    rng = np.random.default_rng(2)
    for rev in (False, True):
        c0, c1, v = sorted_coo(rng, 3000, 200, 200)
        w = random_records(rng, 2500, 200, 200)
        s, n = oracle.wspace_sort(w, rev=rev)
        m0, m1, mv = oracle.wspace_merge(c0, c1, v, s[:n], rev=rev)
        e0, e1, ev = dict_merge(c0, c1, v, w)
        np.testing.assert_array_equal(m0, e0)
        np.testing.assert_array_equal(m1, e1)
        # one addition per shared coordinate: exact in either order
        np.testing.assert_array_equal(mv, ev)


def test_merge_empty_sides():
    rng = np.random.default_rng(3)
    w = random_records(rng, 50, 20, 20)
    s, n = oracle.wspace_sort(w, rev=False)
    e = np.zeros(0, np.int32)
    m0, m1, mv = oracle.wspace_merge(e, e, np.zeros(0), s, rev=False)
    np.testing.assert_array_equal(m0, s["crd"][:, 0])
    np.testing.assert_array_equal(mv, s["val"])
    c0, c1, v = sorted_coo(rng, 40, 20, 20)
    m0, m1, mv = oracle.wspace_merge(c0, c1, v, s[:0], rev=True)
    np.testing.assert_array_equal(m1, c1)
    np.testing.assert_array_equal(mv, v)
