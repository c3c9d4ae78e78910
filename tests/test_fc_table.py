"""Seeded colour tables of the enumerated false-colour subjects.

The reference (EnumeratedFCSubject, FCSubject.h:153-167) draws `size`
distinct colours in [1, 0xFFFFFF] into a std::set and shuffles them with an
unseeded generator; this build draws and shuffles from a seeded splitmix64
so an FC image is reproducible.  The product (bitmap of drawn colours) and
the oracle (std::set, as the reference) build the table independently and
must agree bit for bit; the table is pinned by what the reference's
construction guarantees: distinct colours in range, the ascending set
permuted, every permutation equally likely (Fisher-Yates: 6000 shuffles of
3 elements hit each of the 6 orders 1000 +- 15 %), a deterministic table per
(seed, subject), and the paint semantics of the reference's own tests
(FCSubject.cpp:96-190): ids below the category size get non-black colours,
others black.
"""
import ctypes as C
import itertools

import numpy as np
import pytest


def _oracle(ref, seed, sub, n):
    out = np.zeros(max(1, n), np.uint32)
    rc = ref.lib().ref_fc_color_table(seed, sub, n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert rc == 0
    return out[:n]


@pytest.mark.parametrize("seed,sub,n", [(1, 1, 10), (7, 3, 1000), (123, 11, 50000), (0, 2, 1), (99, 10, 0), (5, 1, 200000)])
def test_product_equals_oracle(pkg, ref, seed, sub, n):
    a = pkg.fc_color_table(seed, sub, n)
    b = _oracle(ref, seed, sub, n)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [1, 17, 4096, 100000])
def test_table_properties(pkg, n):
    t = pkg.fc_color_table(2024, 1, n)
    assert t.size == n and t.min() >= 1 and t.max() <= 0xFFFFFF
    assert np.unique(t).size == n                              # distinct
    if n >= 17:
        assert not np.all(np.diff(t.astype(np.int64)) > 0)     # shuffled
    if n >= 4096:
        assert abs(t.mean() / 0x800000 - 1.0) < 0.03            # uniform draws
    assert np.array_equal(t, pkg.fc_color_table(2024, 1, n))     # deterministic
    assert not np.array_equal(t, pkg.fc_color_table(2025, 1, n)) or n == 1
    assert not np.array_equal(t, pkg.fc_color_table(2024, 2, n)) or n == 1


def test_shuffle_is_uniform(pkg):
    counts = {}
    for seed in range(6000):
        t = pkg.fc_color_table(seed, 3, 3)
        order = tuple(np.argsort(t))
        counts[order] = counts.get(order, 0) + 1
    assert set(counts) == set(itertools.permutations(range(3)))
    assert all(abs(c - 1000) < 150 for c in counts.values()), counts


def test_too_many_ids_refused(pkg):
    with pytest.raises(RuntimeError, match="-2"):
        pkg.fc_color_table(1, 1, 0x1000000)


def test_paint_semantics_of_reference_tests(pkg):
    """GeometryIDFC / PrimitiveIDFC with 3 and 4 categories: ids below the
    size paint distinct non-black colours, ids at or above it black."""
    for size in (3, 4):
        t = pkg.fc_color_table(31, 1, size)
        paint = lambda i: (t[i] if i < size else 0)   # noqa: E731
        cols = [paint(i) for i in range(size + 2)]
        assert all(c != 0 for c in cols[:size]) and cols[size] == 0 and cols[size + 1] == 0
        assert len(set(cols[:size])) == size
