"""GPU tests of the device radix sort (jbgpu_prim_sort): the sort behind the broadphase's Morton order, the pair
cache's sorted keys, the dispatcher's bins and the colour order.  Checked against numpy's stable argsort of the
keys' bit range: the keys and values must come back in exactly that order (a stable sort leaves equal digits in
their input order), at sizes on both sides of every tile size (1024, 2048 and 4096 keys per tile) and pass
count, with uniform, skewed and constant digits."""
import numpy as np
import pytest

import jbgpu

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 255, 256, 1023, 1024, 1025, 4095, 4096, 4097, 70_001, 262_143, 262_144, 524_287, 524_288, 1_100_003]


def _expected(keys, values, lo, hi):
    mask = (1 << (hi - lo)) - 1 if hi - lo < 64 else (1 << 64) - 1
    digits = (keys.astype(np.uint64) >> np.uint64(lo)) & np.uint64(mask)
    order = np.argsort(digits, kind="stable")
    return keys[order], values[order]


@pytest.fixture(scope="module")
def ctx():
    c = jbgpu.Context()
    yield c
    c.close()


@pytest.mark.parametrize("n", SIZES)
def test_sort_pairs_u64_uniform(ctx, n):
    """Random 64-bit keys sorted by bits [3, 45) (six digit passes, the last one narrowed) with their indices."""
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 63, size=n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.int32)
    gk, gv = ctx.prim_sort(keys, vals, 3, 45)
    ek, ev = _expected(keys, vals, 3, 45)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev), f"n={n}"


@pytest.mark.parametrize("n", [4097, 262_144, 1_100_003])
def test_sort_keys_u64_full_width(ctx, n):
    """All 64 bits, keys only (eight passes): the keys come back in ascending unsigned order."""
    rng = np.random.default_rng(7 + n)
    keys = rng.integers(0, 1 << 63, size=n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=n, dtype=np.uint64)
    gk, _ = ctx.prim_sort(keys)
    assert np.array_equal(gk, np.sort(keys))


@pytest.mark.parametrize("n", [5000, 600_000])
def test_sort_skewed_and_constant_digits(ctx, n):
    """Keys whose digits are almost all equal (one digit value holds 99% of a pass) and keys that are all the same:
    every key lands in one bucket, the look-back sums tiles that hold a single digit, and the order is the input
    order."""
    rng = np.random.default_rng(n)
    keys = np.full(n, 0x1234_5678_9abc, dtype=np.uint64)
    flip = rng.random(n) < 0.01
    keys[flip] = rng.integers(0, 1 << 48, size=int(flip.sum()), dtype=np.uint64)
    vals = np.arange(n, dtype=np.int32)
    gk, gv = ctx.prim_sort(keys, vals, 0, 48)
    ek, ev = _expected(keys, vals, 0, 48)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)
    same = np.full(n, 77, dtype=np.uint64)
    gk, gv = ctx.prim_sort(same, vals, 0, 40)
    assert np.array_equal(gk, same) and np.array_equal(gv, vals)


@pytest.mark.parametrize("n", [3000, 300_000, 1_100_003])
def test_sort_pairs_u32(ctx, n):
    """32-bit keys with values, by bits [20, 30) as the colour order sorts its slot keys, and by all 32 bits."""
    rng = np.random.default_rng(3 * n)
    keys = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    vals = rng.integers(-1000, 1000, size=n, dtype=np.int32)
    for lo, hi in ((20, 30), (0, 32)):
        gk, gv = ctx.prim_sort(keys, vals, lo, hi)
        ek, ev = _expected(keys.astype(np.uint64), vals, lo, hi)
        assert np.array_equal(gk, ek.astype(np.uint32)) and np.array_equal(gv, ev), (n, lo, hi)


def test_sort_argument_errors(ctx):
    """Bad key sizes and bit ranges are refused before anything reaches the device; empty input is a no-op."""
    k = np.zeros(10, np.uint64)
    with pytest.raises(jbgpu.JbgpuError):
        ctx.prim_sort(k, None, 10, 5)
    with pytest.raises(jbgpu.JbgpuError):
        ctx.prim_sort(k, None, 0, 65)
    with pytest.raises(ValueError):
        ctx.prim_sort(np.zeros(4, np.float32))
    gk, gv = ctx.prim_sort(np.zeros(0, np.uint64), np.zeros(0, np.int32))
    assert gk.size == 0 and gv.size == 0
