"""GPU parity of the block operations (init, scale, add, norm2, slice,
insert): the drop-in symbols against the reference's own outputs, the native
device entry points against the oracle.  Bit-exact everywhere except the
squared norm, whose summation order differs from the reference's (relative
tolerance 1e-13, stated per test)."""
import math
import os

import numpy as np
import pytest

from checker import Oracle

from aces4_amd import sipgpu as sg

pytestmark = pytest.mark.gpu
o = Oracle()
HERE = os.path.dirname(os.path.abspath(__file__))
Z = np.load(os.path.join(HERE, "golden", "reference_blockops.npz"), allow_pickle=False)
NORM_RTOL = 1e-13


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _keys(prefix):
    return sorted({k.split("_")[0] for k in Z.files if k.startswith(prefix)})


def _shape(i):
    key = "shape%03d" % i
    rank = int(Z[key + "_rank"][0])
    dims = [int(x) for x in Z[key + "_dims"]][:max(rank, 0)]
    return rank, dims, Z[key + "_x"], Z[key + "_y"]


@pytest.mark.parametrize("key", _keys("init") + _keys("scale") + _keys("norm2") + _keys("add"))
def test_dropin_elementwise_golden(key):
    op = key[:-3]
    rank, dims, x, y = _shape(int(Z[key + "_shape"][0]))
    arg = float(Z[key + "_arg"][0])
    want = Z[key + "_out"]
    print(key, op, rank, dims, arg, flush=True)
    r = rank if rank < 0 else None
    if op == "init":
        got, ierr = sg.tensor_block_init(dims, x, arg, rank=r)
    elif op == "scale":
        got, ierr = sg.tensor_block_scale(dims, x, arg, rank=r)
    elif op == "add":
        got, ierr = sg.tensor_block_add(dims, x, y, arg, rank=r)
    else:
        v, ierr = sg.tensor_block_norm2(dims, x, rank=r)
        assert ierr == int(Z[key + "_ierr"][0])
        assert abs(v - want[0]) <= NORM_RTOL * abs(want[0])
        return
    assert ierr == int(Z[key + "_ierr"][0])
    assert np.array_equal(_bits(got), _bits(want))


def _slice_case(key):
    rank = int(Z[key + "_rank"][0])
    r = max(rank, 0)
    te = [int(v) for v in Z[key + "_tens_ext"]][:r]
    se = [int(v) for v in Z[key + "_slice_ext"]][:r]
    eb = [int(v) for v in Z[key + "_ext_beg"]][:r]
    tens = 1.0 + np.arange(max(int(np.prod(te)) if te else 1, 1), dtype=np.float64)
    slc = -1.0 - np.arange(max(int(np.prod(se)) if se else 1, 1), dtype=np.float64)

    def decode(v):
        return np.where(np.asarray(v) == 0, np.nan, np.asarray(v, dtype=np.float64))

    return (rank, te, se, eb, tens, slc, decode(Z[key + "_slice_out"]),
            decode(Z[key + "_insert_out"]), [int(v) for v in Z[key + "_ierr"]])


@pytest.mark.parametrize("key", _keys("slc"))
def test_dropin_slice_insert_golden(key):
    rank, te, se, eb, tens, slc, want_s, want_i, ierrs = _slice_case(key)
    print(key, rank, te, se, eb, flush=True)
    r = rank if rank < 0 else None
    got_s, e1 = sg.tensor_block_slice(tens, te, se, eb, rank=r)
    got_i, e2 = sg.tensor_block_insert(tens, te, slc, se, eb, rank=r)
    assert [e1, e2] == ierrs
    if rank < 0:  # nothing is written
        assert np.isnan(got_s).all() and np.array_equal(got_i, tens)
        return
    np.testing.assert_array_equal(got_s, want_s)
    np.testing.assert_array_equal(got_i, want_i)


def test_dropin_errors_and_empty_slices():
    tens = np.arange(1.0, 13.0)
    out, e = sg.tensor_block_slice(tens, [3, 4], [0, 2], [0, 0], out=np.full(1, 5.0))
    assert e == -1 and out[0] == 5.0
    got, e = sg.tensor_block_insert(tens, [3, 4], np.ones(1), [2, 0], [0, 0])
    assert e == -1 and np.array_equal(got, tens)
    # a slice leaving the block is refused before any device access
    _, e = sg.tensor_block_slice(tens, [3, 4], [2, 2], [2, 0])
    assert e == -107
    got, e = sg.tensor_block_insert(tens, [3, 4], np.ones(4), [2, 2], [0, 3])
    assert e == -107 and np.array_equal(got, tens)
    # degenerate extents: nothing to do, success
    got, e = sg.tensor_block_scale([4, 0], np.full(1, 3.0), 2.0)
    assert e == 0 and got[0] == 3.0
    v, e = sg.tensor_block_norm2([0], np.zeros(1))
    assert e == 0 and v == 0.0


def _slice_shapes(seed, count, max_ext, limit):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        rank = int(rng.integers(1, 7))
        te = [int(v) for v in rng.integers(1, max_ext + 1, size=rank)]
        if np.prod(te) > limit:
            continue
        se = [int(rng.integers(1, t + 1)) for t in te]
        if rng.random() < 0.3:
            k = int(rng.integers(1, rank + 1))
            se[:k] = te[:k]  # leading dims whole: the plan merges them
        eb = [int(rng.integers(0, t - n + 1)) for t, n in zip(te, se)]
        out.append((te, se, eb))
    return out


def test_native_slice_insert_random_vs_oracle():
    rng = np.random.default_rng(11)
    shapes = _slice_shapes(3, 120, 40, 3_000_000)
    shapes += [([30, 15, 30, 15], [15, 15, 10, 15], [15, 0, 20, 0]),
               ([450, 450], [200, 450], [250, 0]), ([64, 1, 129], [1, 1, 129], [63, 0, 0]),
               ([100, 100, 100], [1, 100, 1], [0, 0, 99]), ([7], [7], [0]), ([9], [1], [8])]
    for te, se, eb in shapes:
        print("slice", te, se, eb, flush=True)
        tens = rng.standard_normal(int(np.prod(te)))
        slc = rng.standard_normal(int(np.prod(se)))
        want_s, _ = o.block_slice(tens, te, se, eb)
        want_i, _ = o.block_insert(tens, te, slc, se, eb)
        dt = sg.DeviceArray.from_host(tens)
        ds = sg.DeviceArray(slc.size)
        sg.block_slice(dt.ptr, te, ds.ptr, se, eb)
        assert np.array_equal(_bits(ds.download()), _bits(want_s)), (te, se, eb)
        ds.upload(slc)
        sg.block_insert(dt.ptr, te, ds.ptr, se, eb)
        assert np.array_equal(_bits(dt.download()), _bits(want_i)), (te, se, eb)
        dt.free()
        ds.free()


def test_native_slice_round_trip_and_scalar():
    rng = np.random.default_rng(2)
    te, se, eb = [40, 30, 20, 10], [13, 30, 7, 4], [5, 0, 11, 6]
    tens = rng.standard_normal(int(np.prod(te)))
    dt = sg.DeviceArray.from_host(tens)
    ds = sg.DeviceArray(int(np.prod(se)))
    s = sg.Stream()
    sg.block_slice(dt.ptr, te, ds.ptr, se, eb, stream=s)
    sg.block_init(dt.ptr, te, 0.0, stream=s)
    sg.block_insert(dt.ptr, te, ds.ptr, se, eb, stream=s)
    s.synchronize()
    got = dt.download().reshape(te, order="F")
    want = np.zeros(te)
    region = tuple(slice(b, b + n) for b, n in zip(eb, se))
    want[region] = tens.reshape(te, order="F")[region]
    assert np.array_equal(_bits(got.ravel(order="F")), _bits(want.ravel(order="F")))
    s.destroy()
    # rank 0: the scalar is copied
    a = sg.DeviceArray.from_host(np.array([4.25]))
    b = sg.DeviceArray.from_host(np.array([0.0]))
    lib = sg.lib()
    sg._check(lib.sipgpu_block_slice(None, 0, a.ptr, None, b.ptr, None, None))
    assert b.download()[0] == 4.25
    # out of bounds and empty slices
    with pytest.raises(sg.SipGpuError) as ei:
        sg.block_slice(dt.ptr, te, ds.ptr, [13, 30, 7, 5], [5, 0, 11, 6])
    assert ei.value.code == -107
    sg.block_slice(dt.ptr, te, ds.ptr, [13, 0, 7, 4], eb)  # empty: no-op


@pytest.mark.parametrize("n", [1, 2, 255, 4097, 1_000_003, 6_400_000])
@pytest.mark.parametrize("misalign", [0, 1])
def test_native_elementwise_vs_oracle(n, misalign):
    """Both the 16-byte and the 8-byte access forms (a pointer one double off
    16-byte alignment), odd and even sizes; bit-exact."""
    print("elementwise", n, misalign, flush=True)
    rng = np.random.default_rng(n + misalign)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    dims = [n]
    da, db = sg.DeviceArray(n + 1), sg.DeviceArray(n + 1)
    da.upload(np.full(n + 1, 123.0))  # the element outside the block is a sentinel
    pa, pb = da.at(misalign), db.at(misalign)

    def load():
        da.upload_range(misalign, x)
        db.upload_range(misalign, y)

    def result():
        return da.download_range(misalign, n)

    load()
    sg.block_scale(pa, dims, -0.3)
    assert np.array_equal(_bits(result()), _bits(o.block_scale(dims, x, -0.3)[0]))
    for s in (0.7, 1.0 + 5e-14, 1.0):
        load()
        sg.block_add(pa, pb, dims, s)
        assert np.array_equal(_bits(result()), _bits(o.block_add(dims, x, y, s)[0])), s
    sg.block_init(pa, dims, 1.5)
    assert np.array_equal(_bits(result()), _bits(np.full(n, 1.5)))
    load()
    r = sg.DeviceArray(1)
    sg.block_norm2(pa, dims, r.ptr)
    v = r.download()[0]
    exact = math.fsum((x * x).tolist())
    assert abs(v - exact) <= NORM_RTOL * exact
    sg.block_norm2(pa, dims, r.ptr)
    assert r.download()[0] == v  # fixed order: the same bits again
    # the element next to the block is untouched
    assert da.download_range(n if misalign == 0 else 0, 1)[0] == 123.0
    for d in (da, db, r):
        d.free()


def test_native_scalar_add_always_scales():
    """Rank 0 multiplies by the factor even inside the 1e-13 window
    (tensor_dil_omp.F90:417-418)."""
    a = sg.DeviceArray.from_host(np.array([1.0]))
    b = sg.DeviceArray.from_host(np.array([3.0]))
    s = 1.0 + 5e-14
    sg.block_add(a.ptr, b.ptr, [], s)
    want, _ = o.block_add([], np.array([1.0]), np.array([3.0]), s)
    assert np.array_equal(_bits(a.download()), _bits(want))
    assert a.download()[0] != 4.0


@pytest.mark.parametrize("nparts,n,misalign", [(1, 1000, 0), (5, 1_000_001, 0), (8, 4097, 1),
                                               (3, 3, 1)])
def test_native_add_many_vs_oracle(nparts, n, misalign):
    """t = ((t + p0) + p1) + ...: the oracle's tensor_block_add with factor one,
    once per source in order; bit-exact."""
    print("add_many", nparts, n, misalign, flush=True)
    rng = np.random.default_rng(nparts * 7 + n)
    t = rng.standard_normal(n)
    ps = [rng.standard_normal(n) for _ in range(nparts)]
    want = t.copy()
    for p in ps:
        want, e = o.block_add([n], want, p, 1.0)
        assert e == 0
    dt = sg.DeviceArray(n + 1)
    dt.upload_range(misalign, t)
    dps = []
    for p in ps:
        d = sg.DeviceArray(n + 1)
        d.upload_range(misalign, p)
        dps.append(d)
    sg.block_add_many(dt.at(misalign), [d.at(misalign) for d in dps], [n])
    assert np.array_equal(_bits(dt.download_range(misalign, n)), _bits(want))
    with pytest.raises(sg.SipGpuError):
        sg.block_add_many(dt.ptr, [dps[0].ptr] * 9, [n])
    for d in [dt] + dps:
        d.free()


def test_amplitude_update_sequence_bitwise():
    """The T2 update of rccsd_rhf.sialx (UPDATET2, :2053-2057) as device block
    operations -- Taibj = T2new(b,j,a,i); Taibj += T2new(a,i,b,j); Taibj *=
    0.5; energy_denominator_rhf -- against the same sequence in the oracle,
    bit for bit; then the DIIS error T1 -= Taibj; T1 *= -0.5 (:2030-2031)."""
    rng = np.random.default_rng(17)
    dims = [30, 15, 30, 15]  # (a, i, b, j) of one ring block
    n = int(np.prod(dims))
    t2new = rng.standard_normal(n)
    t1 = rng.standard_normal(n)
    signs = [-1, 1, -1, 1]
    eps = [rng.uniform(0.1, 3, d) if s < 0 else rng.uniform(-2, -0.3, d) for d, s in zip(dims, signs)]
    # oracle: (a,i,b,j) -> (b,j,a,i) is dim_transp [1, 3, 4, 1, 2]
    tr = [1, 3, 4, 1, 2]
    want, e = o.copy(dims, tr, t2new)
    assert e == 0
    want, e = o.block_add(dims, want, t2new, 1.0)
    want, e = o.block_scale(dims, want, 0.5)
    want = o.energy_denominator(dims, want, eps, signs)
    diis, e = o.block_add(dims, t1, want, -1.0)
    diis, e = o.block_scale(dims, diis, -0.5)
    # device
    d_new = sg.DeviceArray.from_host(t2new)
    d_t = sg.DeviceArray(n)
    d_t1 = sg.DeviceArray.from_host(t1)
    ed = [sg.DeviceArray.from_host(x) for x in eps]
    sg.permute(dims, tr, d_new.ptr, d_t.ptr)
    sg.block_add(d_t.ptr, d_new.ptr, dims, 1.0)
    sg.block_scale(d_t.ptr, dims, 0.5)
    sg.energy_denominator(dims, d_t.ptr, [x.ptr for x in ed], signs)
    sg.block_add(d_t1.ptr, d_t.ptr, dims, -1.0)
    sg.block_scale(d_t1.ptr, dims, -0.5)
    sg.synchronize()
    assert np.array_equal(_bits(d_t.download()), _bits(want))
    assert np.array_equal(_bits(d_t1.download()), _bits(diis))


def test_dropin_slice_insert_rank_ten():
    """A rank-10 slice (the reference module takes any rank) through the
    drop-in symbols, bit-exact against the oracle."""
    te = [3, 1, 4, 1, 5, 2, 1, 3, 2, 2]
    se = [3, 1, 2, 1, 5, 2, 1, 1, 2, 1]
    eb = [0, 0, 1, 0, 0, 0, 0, 2, 0, 1]
    rng = np.random.default_rng(23)
    tens = rng.standard_normal(int(np.prod(te)))
    slc = rng.standard_normal(int(np.prod(se)))
    want_s, e = o.block_slice(tens, te, se, eb)
    assert e == 0
    got_s, e = sg.tensor_block_slice(tens, te, se, eb)
    assert e == 0 and np.array_equal(_bits(got_s), _bits(want_s))
    want_i, _ = o.block_insert(tens, te, slc, se, eb)
    got_i, e = sg.tensor_block_insert(tens, te, slc, se, eb)
    assert e == 0 and np.array_equal(_bits(got_i), _bits(want_i))
