"""The oracle's block operations (init, scale, norm2, add, slice, insert)
pinned three ways: against the reference's own outputs committed in
tests/golden/reference_blockops.npz, against the reference build directly
where it is present, and against closed forms (numpy slicing and elementwise
arithmetic, an exactly rounded sum)."""
import math
import os

import numpy as np
import pytest

from checker import Oracle, RefLib

HERE = os.path.dirname(os.path.abspath(__file__))
Z = np.load(os.path.join(HERE, "golden", "reference_blockops.npz"), allow_pickle=False)
O = Oracle()


def _keys(prefix):
    return sorted({k.split("_")[0] for k in Z.files if k.startswith(prefix)})


def _shape(i):
    key = "shape%03d" % i
    rank = int(Z[key + "_rank"][0])
    dims = [int(x) for x in Z[key + "_dims"]][:max(rank, 0)]
    return rank, dims, Z[key + "_x"], Z[key + "_y"]


def _run(op, rank, dims, x, y, arg):
    if rank < 0:
        # the checker bindings derive the rank from the list length; call the
        # C restatement directly for a negative rank
        import ctypes
        from checker import _dptr, _ints, _iptr
        t = np.array(x, dtype=np.float64)
        e = ctypes.c_int(0)
        d = _ints([0])
        if op == "init":
            O.lib.orc_block_init(rank, _iptr(d), _dptr(t), ctypes.c_double(arg), ctypes.byref(e))
        elif op == "scale":
            O.lib.orc_block_scale(rank, _iptr(d), _dptr(t), ctypes.c_double(arg), ctypes.byref(e))
        elif op == "norm2":
            O.lib.orc_block_norm2.restype = ctypes.c_double
            v = O.lib.orc_block_norm2(rank, _iptr(d), _dptr(t), ctypes.byref(e))
            return np.array([v]), e.value
        else:
            O.lib.orc_tensor_block_add(rank, _iptr(d), _dptr(t), _dptr(np.array(y)),
                                       ctypes.c_double(arg), ctypes.byref(e))
        return t, e.value
    if op == "init":
        return O.block_init(dims, x, arg)
    if op == "scale":
        return O.block_scale(dims, x, arg)
    if op == "norm2":
        v, e = O.block_norm2(dims, x)
        return np.array([v]), e
    return O.block_add(dims, x, y, arg)


@pytest.mark.parametrize("key", _keys("init") + _keys("scale") + _keys("norm2") + _keys("add"))
def test_elementwise_matches_reference_vectors(key):
    op = key[:-3]
    rank, dims, x, y = _shape(int(Z[key + "_shape"][0]))
    arg = float(Z[key + "_arg"][0])
    out, ierr = _run(op, rank, dims, x, y, arg)
    assert ierr == int(Z[key + "_ierr"][0])
    # bit-exact, norm2 included: one reference thread sums in ascending order
    np.testing.assert_array_equal(out.view(np.uint64), Z[key + "_out"].view(np.uint64))


def _slice_case(key):
    rank = int(Z[key + "_rank"][0])
    r = max(rank, 0)
    te = [int(v) for v in Z[key + "_tens_ext"]][:r]
    se = [int(v) for v in Z[key + "_slice_ext"]][:r]
    eb = [int(v) for v in Z[key + "_ext_beg"]][:r]
    nt = max(int(np.prod(te)) if te else 1, 1)
    ns = max(int(np.prod(se)) if se else 1, 1)
    tens = 1.0 + np.arange(nt, dtype=np.float64)
    slc = -1.0 - np.arange(ns, dtype=np.float64)

    def decode(v):
        v = np.asarray(v)
        return np.where(v == 0, np.nan, v.astype(np.float64))

    return rank, te, se, eb, tens, slc, decode(Z[key + "_slice_out"]), decode(
        Z[key + "_insert_out"]), [int(v) for v in Z[key + "_ierr"]]


@pytest.mark.parametrize("key", _keys("slc"))
def test_slice_insert_match_reference_vectors(key):
    rank, te, se, eb, tens, slc, want_slice, want_ins, ierrs = _slice_case(key)
    if rank < 0:
        import ctypes
        from checker import _dptr, _ints, _iptr
        out = np.full(1, np.nan)
        e = ctypes.c_int(0)
        d = _ints([0])
        O.lib.orc_block_slice(rank, _dptr(tens), _iptr(d), _dptr(out), _iptr(d), _iptr(d),
                              ctypes.byref(e))
        assert e.value == ierrs[0] == 1
        t = tens.copy()
        O.lib.orc_block_insert(rank, _dptr(t), _iptr(d), _dptr(slc), _iptr(d), _iptr(d),
                               ctypes.byref(e))
        assert e.value == ierrs[1] == 1
        np.testing.assert_array_equal(t, tens)
        return
    out, e1 = O.block_slice(tens, te, se, eb)
    ins, e2 = O.block_insert(tens, te, slc, se, eb)
    assert [e1, e2] == ierrs
    np.testing.assert_array_equal(out, want_slice)
    np.testing.assert_array_equal(ins, want_ins)


def _np_region(te, se, eb):
    return tuple(slice(b, b + n) for b, n in zip(eb, se))


@pytest.mark.parametrize("seed", range(12))
def test_slice_insert_closed_form(seed):
    """Column-major numpy slicing is the closed form of both operations."""
    rng = np.random.default_rng(seed)
    rank = 1 + seed % 6
    te = [int(v) for v in rng.integers(1, 9, size=rank)]
    se = [int(rng.integers(1, t + 1)) for t in te]
    eb = [int(rng.integers(0, t - n + 1)) for t, n in zip(te, se)]
    tens = rng.standard_normal(int(np.prod(te)))
    slc = rng.standard_normal(int(np.prod(se)))
    T = tens.reshape(te, order="F")
    out, e = O.block_slice(tens, te, se, eb)
    assert e == 0
    np.testing.assert_array_equal(out, T[_np_region(te, se, eb)].ravel(order="F"))
    ins, e = O.block_insert(tens, te, slc, se, eb)
    assert e == 0
    W = T.copy()
    W[_np_region(te, se, eb)] = slc.reshape(se, order="F")
    np.testing.assert_array_equal(ins, W.ravel(order="F"))


def test_slice_insert_round_trip():
    rng = np.random.default_rng(7)
    te, se, eb = [6, 5, 4, 3], [2, 5, 3, 1], [3, 0, 1, 2]
    tens = rng.standard_normal(int(np.prod(te)))
    out, _ = O.block_slice(tens, te, se, eb)
    back, _ = O.block_insert(tens, te, out, se, eb)
    np.testing.assert_array_equal(back, tens)


def test_empty_slice_is_rejected_and_writes_nothing():
    tens = np.arange(12.0)
    out, e = O.block_slice(tens, [3, 4], [0, 2], [0, 0], out=np.full(1, 5.0))
    assert e == -1 and out[0] == 5.0
    ins, e = O.block_insert(tens, [3, 4], np.ones(1), [2, 0], [0, 0])
    assert e == -1
    np.testing.assert_array_equal(ins, tens)


def test_elementwise_closed_forms():
    rng = np.random.default_rng(3)
    dims = [5, 7, 3]
    x, y = rng.standard_normal(105), rng.standard_normal(105)
    np.testing.assert_array_equal(O.block_init(dims, x, -3.5)[0], np.full(105, -3.5))
    np.testing.assert_array_equal(O.block_scale(dims, x, 0.37)[0], x * 0.37)
    np.testing.assert_array_equal(O.block_add(dims, x, y, 0.37)[0], x + y * 0.37)
    # a factor within 1e-13 of one is not applied (tensor_dil_omp.F90:421)
    s = 1.0 + 6e-14
    assert not np.array_equal(x + y * s, x + y)
    np.testing.assert_array_equal(O.block_add(dims, x, y, s)[0], x + y)
    np.testing.assert_array_equal(O.block_add(dims, x, y, 1.0 + 3e-13)[0], x + y * (1.0 + 3e-13))
    v, e = O.block_norm2(dims, x)
    exact = math.fsum(float(t) * float(t) for t in x)  # products exact to 1 ulp
    assert e == 0 and abs(v - exact) <= 1e-14 * exact
    # scalars (rank 0) and degenerate extents
    assert O.block_norm2([], np.array([-3.0]))[0] == 9.0
    assert O.block_scale([], np.array([2.0]), 4.0)[0][0] == 8.0
    t, e = O.block_init([2, 0, 3], np.full(1, 7.0), 1.0)
    assert e == 0 and t[0] == 7.0


@pytest.mark.skipif(not RefLib.available(), reason="reference build not present")
@pytest.mark.parametrize("seed", range(8))
def test_oracle_matches_reference_build(seed):
    ref = RefLib()
    rng = np.random.default_rng(100 + seed)
    rank = 1 + seed % 6
    te = [int(v) for v in rng.integers(1, 9, size=rank)]
    se = [int(rng.integers(1, t + 1)) for t in te]
    eb = [int(rng.integers(0, t - n + 1)) for t, n in zip(te, se)]
    tens = rng.standard_normal(int(np.prod(te)))
    slc = rng.standard_normal(int(np.prod(se)))
    a, e1 = O.block_slice(tens, te, se, eb)
    b, e2 = ref.block_slice(tens, te, se, eb, nthreads=8)
    assert e1 == e2 == 0
    np.testing.assert_array_equal(a, b)
    a, e1 = O.block_insert(tens, te, slc, se, eb)
    b, e2 = ref.block_insert(tens, te, slc, se, eb, nthreads=8)
    assert e1 == e2 == 0
    np.testing.assert_array_equal(a, b)
    x, y = tens, rng.standard_normal(tens.size)
    for s in (0.25, 1.0 + 2e-14, -1.5):
        a, _ = O.block_add(te, x, y, s)
        b = x.copy()
        assert ref.add(te, b, y, s) == 0
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(O.block_scale(te, x, s)[0], ref.block_scale(te, x, s)[0])
    np.testing.assert_array_equal(O.block_init(te, x, 0.5)[0], ref.block_init(te, x, 0.5)[0])
    assert O.block_norm2(te, x)[0] == ref.block_norm2(te, x, nthreads=1)[0]
