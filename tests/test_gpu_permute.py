"""GPU parity of the permute super-instruction (bit-exact against the
reference's golden vectors and the oracle)."""
import hashlib

import numpy as np
import pytest

import golden_io as g
from checker import Oracle

from aces4_amd import sipgpu as sg
from aces4_amd.workloads import uniform

pytestmark = pytest.mark.gpu
o = Oracle()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("key", g.keys("perm"))
def test_dropin_permute_golden(key):
    dims, transp, x, y = g.permute_case(key)
    out, ierr = sg.tensor_block_copy(dims, transp, x)
    assert ierr == 0
    assert np.array_equal(_bits(out), _bits(y))


@pytest.mark.parametrize("case", g.manifest()["hashed_permutes"], ids=lambda c: str(c["dims"]))
def test_dropin_permute_golden_hashed(case):
    x = uniform(case["seed"], int(np.prod(case["dims"])), stream=0, lo=-1.0, hi=1.0)
    out, ierr = sg.tensor_block_copy(case["dims"], case["transp"], x)
    assert ierr == 0
    assert hashlib.sha256(out.tobytes()).hexdigest() == case["out_sha256"]


def _random_perm_cases(seed, n, max_ext):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        rank = int(rng.integers(1, 7))
        while True:
            dims = [int(x) for x in rng.integers(1, max_ext + 1, size=rank)]
            if np.prod(dims) <= 3_000_000:
                break
        perm = [int(x) + 1 for x in rng.permutation(rank)]
        yield dims, [1] + perm


def test_native_permute_random_vs_oracle():
    rng = np.random.default_rng(5)
    for dims, transp in _random_perm_cases(31, 150, 40):
        print("permute", dims, transp, flush=True)
        x = rng.standard_normal(int(np.prod(dims)))
        want, _ = o.copy(dims, transp, x)
        din = sg.DeviceArray.from_host(x)
        dout = sg.DeviceArray(x.size)
        sg.permute(dims, transp, din.ptr, dout.ptr)
        got = dout.download()
        assert np.array_equal(_bits(got), _bits(want)), (dims, transp)
        din.free()
        dout.free()


@pytest.mark.parametrize("dims,perm", [
    ([30, 15, 30, 15], [3, 4, 1, 2]),          # ring block, (ai)<->(bj)
    ([64, 64, 64, 64], [4, 3, 2, 1]),          # 16.7 M doubles, full reversal
    ([40, 40, 40, 40], [2, 1, 3, 4]),          # only the two fastest swap
    ([20, 20, 20, 20, 20, 20], [6, 1, 5, 2, 4, 3]),
    ([3, 2000, 7], [2, 3, 1]),
    ([2, 2, 2, 2, 2, 2], [2, 1, 4, 3, 6, 5]),
    ([1000003], [1]),
])
def test_native_permute_shapes(dims, perm):
    print("permute", dims, perm, flush=True)
    n = int(np.prod(dims))
    x = uniform(77, n, lo=-1.0, hi=1.0)
    want, _ = o.copy(dims, [1] + perm, x)
    din = sg.DeviceArray.from_host(x)
    dout = sg.DeviceArray(n)
    sg.permute(dims, [1] + perm, din.ptr, dout.ptr)
    assert np.array_equal(_bits(dout.download()), _bits(want))


def test_permute_errors():
    x = np.arange(6.0)
    _, ierr = sg.tensor_block_copy([2, 3], [1, 1, 1], x)
    assert ierr == -103
    _, ierr = sg.tensor_block_copy([2, 3], [1, 3, 1], x)
    assert ierr == -103
    # negative rank: ierr = rank, as in the reference
    import ctypes
    ierr = ctypes.c_int(0)
    d = np.array([2], np.int32)
    t = np.array([1, 1], np.int32)
    sg.lib().tensor_block_copy__(ctypes.byref(ctypes.c_int(8)), ctypes.byref(ctypes.c_int(-2)),
                                 d.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 t.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 ctypes.byref(ierr))
    assert ierr.value == -2


def test_scalar_permute():
    out, ierr = sg.tensor_block_copy([], [1], np.array([3.25]))
    assert ierr == 0 and out[0] == 3.25


def _batch_cases():
    """Entries of one permute batch: random ranks 1..6, the shapes above, an
    identity, extent-1 dimensions, a scalar and an empty block."""
    cases = list(_random_perm_cases(43, 40, 24))
    cases += [([30, 15, 30, 15], [1, 3, 4, 1, 2]), ([64, 64, 32, 16], [1, 4, 3, 2, 1]),
              ([7, 5, 3], [1, 1, 2, 3]), ([1, 9, 1, 4], [1, 2, 4, 3, 1]),
              ([3, 2000, 7], [1, 2, 3, 1]), ([], [1]), ([3, 0, 4], [1, 2, 1, 3])]
    return cases


def test_permute_batch_vs_oracle():
    """One launch over 47 permutes of mixed ranks and shapes, each output bit
    for bit the oracle's copy; a second execution gives the same bits."""
    cases = _batch_cases()
    sizes = [int(np.prod(d)) if d else 1 for d, _ in cases]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    x = uniform(91, int(offs[-1]), lo=-1.0, hi=1.0)
    din = sg.DeviceArray.from_host(x)
    dout = sg.DeviceArray(int(offs[-1]))
    entries = [sg.permutation(d, t, din.ptr + 8 * int(offs[i]), dout.ptr + 8 * int(offs[i]))
               for i, (d, t) in enumerate(cases)]
    b = sg.PermuteBatch(entries)
    assert b.elements == int(offs[-1])
    print("permute batch", len(cases), "entries", b.workgroups, "workgroups", flush=True)
    for rep in range(2):
        dout.zero()
        b.execute()
        sg.synchronize()
        got = dout.download()
        for i, (d, t) in enumerate(cases):
            seg = x[offs[i]:offs[i + 1]]
            want = seg.copy() if not d or seg.size == 0 else o.copy(d, t, seg)[0]
            assert np.array_equal(_bits(got[offs[i]:offs[i + 1]]), _bits(want)), (rep, d, t)
    b.destroy()
    din.free()
    dout.free()


def test_permute_batch_errors():
    din, dout = sg.DeviceArray(6), sg.DeviceArray(6)
    good = sg.permutation([2, 3], [1, 2, 1], din.ptr, dout.ptr)
    with pytest.raises(sg.SipGpuError) as e:
        sg.PermuteBatch([good, sg.permutation([2, 3], [1, 1, 1], din.ptr, dout.ptr)])
    assert e.value.code == -103
    with pytest.raises(sg.SipGpuError) as e:
        sg.PermuteBatch([sg.permutation([2, 3], [1, 2, 1], 0, dout.ptr)])
    assert e.value.code == -107  # SIPGPU_EARG
    empty = sg.PermuteBatch([])
    assert empty.elements == 0 and empty.workgroups == 0
    empty.execute()
    empty.destroy()
    din.free()
    dout.free()
