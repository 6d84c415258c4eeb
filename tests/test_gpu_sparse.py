"""GPU parity on the block-sparse ladder (uneven segments, irrep-selected
blocks, data generated in HBM) and of the device data generator.

Operand blocks are read back from the device and contracted by the oracle;
FP64 batch within 1e-13, split path within 1e-12 relative Frobenius per
result block.
"""
import numpy as np
import pytest

from cases import rel_fro
from checker import Oracle

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu
o = Oracle()


def small():
    return wl.sparse_ladder(o=16, v=48, seg_lo=4, seg_hi=12, seed=3)


def oracle_block(w, rb, Vh, Th):
    V, T, R, ptrn = w["L"], w["R"], w["D"], w["ptrn"]
    out = np.zeros(int(np.prod(R.dims(rb))))
    for db, vb, tb in w["entries"]:
        if db != rb:
            continue
        tmp, e = o.contract(ptrn, V.download_block(vb, Vh), V.dims(vb),
                            T.download_block(tb, Th), T.dims(tb), R.dims(rb))
        assert e == 0
        out += tmp
    return out


@pytest.mark.parametrize("moduli,tol", [(None, 1e-13), (14, 1e-12)], ids=["fp64", "split14"])
def test_sparse_ladder_vs_oracle(moduli, tol):
    w = small()
    print("sparse ladder", len(w["entries"]), moduli, flush=True)
    for k in ("D", "L", "R"):
        w[k].to_device()
    b = wl.batch_for(w, moduli=moduli)
    b.execute()
    sg.synchronize()
    got = w["D"].dev.download()
    Vh, Th = w["L"].dev.download(), w["R"].dev.download()
    assert abs(b.flops - wl.flops(w)) <= 1e-9 * wl.flops(w)
    worst = 0.0
    for rb in w["D"].block_ids():
        worst = max(worst, rel_fro(w["D"].download_block(rb, got), oracle_block(w, rb, Vh, Th)))
    assert worst <= tol


def test_fill_random_matches_host_generator():
    n = 100003
    a = sg.DeviceArray(n)
    sg.fill_random(a.ptr, n, 7, 5, "uniform", -0.5, 0.5)
    sg.synchronize()
    assert np.array_equal(a.download(), wl.uniform(7, n, stream=5, lo=-0.5, hi=0.5))
    sg.fill_random(a.ptr, n, 7, 5, "normal", 0.1)
    sg.synchronize()
    host = wl.normal(7, n, stream=5, sigma=0.1)
    got = a.download()
    assert np.max(np.abs(got - host)) <= 1e-14 * np.max(np.abs(host))
    a.free()


def full_size_sample(w, blocks):
    """Oracle results of `blocks` result blocks of the full-size ladder, each
    the sum over its pairs of the oracle's contraction (operand blocks read
    back from the device, pairs in parallel: the oracle's C calls release
    the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    V, T, R, ptrn = w["L"], w["R"], w["D"], w["ptrn"]
    want = {}
    for rb in blocks:
        pairs = [(vb, tb) for db, vb, tb in w["entries"] if db == rb]

        def one(p):
            vb, tb = p
            tmp, e = o.contract(ptrn, V.download_block(vb), V.dims(vb), T.download_block(tb),
                                T.dims(tb), R.dims(rb))
            assert e == 0
            return tmp
        with ThreadPoolExecutor(8) as ex:
            tmps = list(ex.map(one, pairs))
        acc = np.zeros(int(np.prod(R.dims(rb))))
        for t in tmps:  # summed in pair order, as the SIP's R += TMP sequence
            acc += t
        want[rb] = acc
    return want


def test_sparse_ladder_full_size_sampled_blocks():
    """The block-sparse ladder at its full size (o = 40, v = 200, segments
    12..40, 22500 pairs): FP64 and split (14 moduli) batches, three result
    blocks (seeded choice) against the oracle: 1e-13 and 1e-12."""
    w = wl.sparse_ladder()
    for k in ("D", "L", "R"):
        w[k].to_device()
    R = w["D"]
    rng = np.random.default_rng(11)
    ids = list(R.block_ids())
    blocks = [ids[int(i)] for i in rng.choice(len(ids), size=3, replace=False)]
    print("sparse ladder full size, blocks", blocks, flush=True)
    want = full_size_sample(w, blocks)
    for moduli, tol in ((None, 1e-13), (14, 1e-12)):
        R.dev.zero()
        b = wl.batch_for(w, moduli=moduli)
        b.execute()
        sg.synchronize()
        for rb in blocks:
            assert rel_fro(R.download_block(rb), want[rb]) <= tol, (moduli, rb)
        b.destroy()
    for k in ("D", "L", "R"):
        w[k].dev.free()
