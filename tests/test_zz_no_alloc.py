"""GPU tests of pbf_sort_by_key_u32_noalloc, the sort_by_key variant that
takes its work space from a fixed device array instead of allocating.

Each call is checked against pbf_sort_by_key_u32 (the allocating entry
point) on the same input, and both against numpy's stable argsort: bit for
bit, keys and values.  Calls run back to back with shrinking inputs, so a
later call must not see what an earlier one left in the shared work space.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, key_bits, seed):
    rng = np.random.default_rng(seed)
    # keys over the whole key_bits range, drawn from a pool of n/8 values:
    # runs of equal keys whose values must keep input order
    pool = rng.integers(0, 1 << key_bits, size=max(1, n // 8), dtype=np.uint64)
    keys = pool[rng.integers(0, pool.size, size=n)].astype(np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    return keys, vals


def _expected(keys, vals):
    order = np.argsort(keys, kind="stable")
    return keys[order], vals[order]


def test_noalloc_sort_equals_allocating_sort_back_to_back(pbf):
    # (n, key bits): 3 radix passes leave the result in the work space (copied
    # back), 2 passes leave it in place; each call smaller than the one before
    cases = [(pbf.NOALLOC_MAX_PAIRS, 20), (300_001, 24), (70_000, 13), (1_025, 16), (2, 8), (1, 32), (0, 32)]
    for i, (n, bits) in enumerate(cases):
        print(f"noalloc sort: call {i}, n={n}, key_bits={bits}", flush=True)
        keys, vals = _case(n, bits, seed=100 + i)
        ek, ev = _expected(keys, vals)
        nk, nv = pbf.sort_by_key_u32(keys, vals, bits, no_alloc=True)
        ak, av = pbf.sort_by_key_u32(keys, vals, bits)
        assert np.array_equal(nk, ak) and np.array_equal(nv, av), f"call {i}: differs from pbf_sort_by_key_u32"
        assert np.array_equal(nk, ek) and np.array_equal(nv, ev), f"call {i}: not the stable order"


def test_noalloc_sort_repeated_calls_same_input(pbf):
    """Two calls in a row on one input, then a smaller input: same results as the allocating sort."""
    keys, vals = _case(500_000, 22, seed=7)
    ak, av = pbf.sort_by_key_u32(keys, vals, 22)
    for _ in range(2):
        nk, nv = pbf.sort_by_key_u32(keys, vals, 22, no_alloc=True)
        assert np.array_equal(nk, ak) and np.array_equal(nv, av)
    small_k, small_v = keys[:4_097].copy(), vals[:4_097].copy()
    sk, sv = pbf.sort_by_key_u32(small_k, small_v, 22)
    nk, nv = pbf.sort_by_key_u32(small_k, small_v, 22, no_alloc=True)
    assert np.array_equal(nk, sk) and np.array_equal(nv, sv)


def test_noalloc_sort_rejects_more_than_its_work_space(pbf):
    """n above PBF_NOALLOC_MAX_PAIRS is refused before any device work."""
    n = pbf.NOALLOC_MAX_PAIRS + 1
    kb, vb = pbf.DeviceBuffer(n * 4), pbf.DeviceBuffer(n * 4)
    rc = pbf.lib().pbf_sort_by_key_u32_noalloc(kb.ptr, vb.ptr, n, 32)
    assert rc == pbf.PBF_ERR_CAPACITY
    assert "PBF_NOALLOC_MAX_PAIRS" in pbf.lib().pbf_last_error().decode()
