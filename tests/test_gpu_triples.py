"""GPU parity of the triples path: W blocks from the twelve permuted
contractions (FP64 batches, and the split path) and the fused
denominator / energy reduction, against the oracle's sequence: every
contraction assigned into a temporary (tensor_block_contract) and added
(tensor_block_add), energy_denominator_rhf, then the full contraction to a
scalar (fcontract)."""
import numpy as np
import pytest

from cases import rel_fro
from checker import Oracle

from aces4_amd import sipgpu as sg
from aces4_amd import triples as tp

pytestmark = pytest.mark.gpu
o = Oracle()


@pytest.fixture(scope="module")
def small():
    w = tp.triples(o_segs=(2, 3), v_segs=(4, 3, 5), seed=5)
    tp.to_device(w)
    host = {k: w[k].dev.download() for k in ("V1", "V2", "T")}
    return w, host


def oracle_tuple(w, host, tup):
    dims = tp.tuple_dims(w, tup)
    W = np.zeros(int(np.prod(dims)))
    for ptrn, pairs in tp.tuple_entries(w, tup):
        for ln, lb, rn, rb in pairs:
            L, R = w[ln], w[rn]
            tmp, e = o.contract(ptrn, L.download_block(lb, host[ln]), L.dims(lb),
                                R.download_block(rb, host[rn]), R.dims(rb), dims)
            assert e == 0
            W = W + tmp
    segs = {"v": w["v_segs"], "o": w["o_segs"]}
    eps = []
    for lab, s in zip(tp.W_LABELS, tup):
        kind = w["kind"][lab]
        base = w["e_vir"] if kind == "v" else w["e_occ"]
        off = tp.seg_offset(segs[kind], s)
        eps.append(base[off:off + segs[kind][s]])
    t = o.energy_denominator(dims, W, eps, tp.W_SIGNS)
    return W, o.fcontract(t, W)


TUPLES = [(0, 1, 2, 0, 1, 0), (1, 1, 1, 1, 1, 1), (2, 0, 1, 1, 0, 1)]


@pytest.mark.parametrize("moduli,wtol,etol", [(None, 1e-13, 1e-13), (14, 1e-12, 1e-12)],
                         ids=["fp64", "split14"])
@pytest.mark.parametrize("tup", TUPLES, ids=[str(t) for t in TUPLES])
def test_tuple_energy_vs_oracle(small, tup, moduli, wtol, etol):
    w, host = small
    print("triples tuple", tup, moduli, flush=True)
    dims = tp.tuple_dims(w, tup)
    wbuf = sg.DeviceArray(int(np.prod(dims)))
    out = sg.DeviceArray(1)
    out.zero()
    te = tp.TupleEnergy(w, tup, wbuf.ptr, out.ptr, moduli=moduli)
    te.execute()
    sg.synchronize()
    W_ref, E_ref = oracle_tuple(w, host, tup)
    assert rel_fro(wbuf.download(), W_ref) <= wtol
    E = out.download()[0]
    assert abs(E - E_ref) <= etol * abs(E_ref)
    # accumulates: a second execution doubles the energy
    te.execute()
    sg.synchronize()
    assert abs(out.download()[0] - 2 * E) <= 1e-14 * abs(E)
    te.destroy()


def test_denominator_kernel_bitwise():
    rng = np.random.default_rng(3)
    dims = [3, 2, 4, 2, 5, 3]
    x = rng.standard_normal(int(np.prod(dims)))
    eps = [rng.uniform(0.1, 3, d) if s < 0 else rng.uniform(-2, -0.3, d)
           for d, s in zip(dims, [-1, 1, -1, 1, -1, 1])]
    xd = sg.DeviceArray.from_host(x)
    ed = [sg.DeviceArray.from_host(e) for e in eps]
    sg.energy_denominator(dims, xd.ptr, [e.ptr for e in ed], [-1, 1, -1, 1, -1, 1])
    sg.synchronize()
    want = o.energy_denominator(dims, x, eps, [-1, 1, -1, 1, -1, 1])
    assert np.array_equal(xd.download(), want)
    out = sg.DeviceArray(1)
    out.zero()
    xd.upload(x)
    sg.energy_sum(dims, xd.ptr, xd.ptr, [e.ptr for e in ed], [-1, 1, -1, 1, -1, 1], 0.5, out.ptr)
    sg.synchronize()
    ref = 0.5 * o.fcontract(want, x)
    assert abs(out.download()[0] - ref) <= 1e-14 * abs(ref)


def test_denominator_kernel_matches_the_reference_loops():
    """The GPU denominator against the reference's do_rhf_den2/4/6 outputs
    (tests/golden/reference_energy_den.npz, non-dyadic energies), bit for bit."""
    from test_oracle_energy import den_cases
    for routine, dims, signs, eps, x, want in den_cases():
        print("denominator fixture", routine, dims, flush=True)
        xd = sg.DeviceArray.from_host(x)
        ed = [sg.DeviceArray.from_host(e) for e in eps]
        sg.energy_denominator(dims, xd.ptr, [e.ptr for e in ed], signs)
        sg.synchronize()
        assert np.array_equal(xd.download(), want), routine
        xd.free()
        for e in ed:
            e.free()


@pytest.mark.parametrize("moduli", [None, 14], ids=["fp64", "split14"])
def test_tuple_with_partial_buffers_is_bitwise_equal(small, moduli):
    """Batches assigned into separate buffers and summed in batch order give
    the same W bits and the same energy as batches accumulated into W, on one
    stream and split over two."""
    w, host = small
    tup = TUPLES[0]
    dims = tp.tuple_dims(w, tup)
    n = int(np.prod(dims))
    print("triples partial buffers", tup, moduli, flush=True)
    results = []
    for use_parts, two in ((False, False), (True, False), (True, True)):
        wbuf = sg.DeviceArray(n)
        out = sg.DeviceArray(1)
        out.zero()
        parts = [sg.DeviceArray(n) for _ in range(5)] if use_parts else None
        te = tp.TupleEnergy(w, tup, wbuf.ptr, out.ptr, moduli=moduli,
                            parts=[p.ptr for p in parts] if parts else None, two_streams=two)
        s = sg.Stream()
        for _ in range(2):  # a second run accumulates a second energy
            te.execute(s)
        s.synchronize()
        results.append((wbuf.download(), out.download()[0]))
        te.destroy()
        s.destroy()
    (w0, e0), *rest = results
    for w1, e1 in rest:
        assert np.array_equal(w0.view(np.uint64), w1.view(np.uint64))
        assert e0 == e1


def test_energy_sum_parts_equals_the_two_passes():
    """sipgpu_energy_sum_parts = block_add_many then energy_sum(x = y = t), bit
    for bit, on an odd-sized rank-6 block."""
    rng = np.random.default_rng(9)
    dims = [3, 4, 5, 2, 3, 7]
    n = int(np.prod(dims))
    signs = [-1, -1, -1, 1, 1, 1]
    eps = [rng.uniform(0.1, 3, d) if s < 0 else rng.uniform(-2, -0.3, d) for d, s in zip(dims, signs)]
    ed = [sg.DeviceArray.from_host(e) for e in eps]
    t0 = rng.standard_normal(n)
    ps = [rng.standard_normal(n) for _ in range(5)]
    dps = [sg.DeviceArray.from_host(p) for p in ps]
    a, b = sg.DeviceArray.from_host(t0), sg.DeviceArray.from_host(t0)
    oa, ob = sg.DeviceArray(1), sg.DeviceArray(1)
    oa.zero()
    ob.zero()
    sg.block_add_many(a.ptr, [d.ptr for d in dps], dims)
    sg.energy_sum(dims, a.ptr, a.ptr, [e.ptr for e in ed], signs, 0.5, oa.ptr)
    sg.energy_sum_parts(dims, b.ptr, [d.ptr for d in dps], [e.ptr for e in ed], signs, 0.5, ob.ptr)
    sg.synchronize()
    assert np.array_equal(a.download().view(np.uint64), b.download().view(np.uint64))
    assert oa.download()[0] == ob.download()[0]
    want = t0.copy()
    for p in ps:
        want = want + p
    assert np.array_equal(b.download(), want)


def test_tuned_tuple_with_parts_vs_oracle(small):
    """Autotuned batches into partial buffers: W and the energy against the
    oracle's contract + add sequence."""
    w, host = small
    tup = TUPLES[2]
    dims = tp.tuple_dims(w, tup)
    n = int(np.prod(dims))
    print("triples tuned with parts", tup, flush=True)
    wbuf = sg.DeviceArray(n)
    out = sg.DeviceArray(1)
    out.zero()
    parts = [sg.DeviceArray(n) for _ in range(5)]
    te = tp.TupleEnergy(w, tup, wbuf.ptr, out.ptr, parts=[p.ptr for p in parts], tune=True)
    te.execute()
    sg.synchronize()
    W_ref, E_ref = oracle_tuple(w, host, tup)
    assert rel_fro(wbuf.download(), W_ref) <= 1e-13
    assert abs(out.download()[0] - E_ref) <= 1e-13 * abs(E_ref)
    te.destroy()
    with pytest.raises(ValueError):
        tp.TupleEnergy(w, tup, wbuf.ptr, out.ptr, parts=[wbuf.ptr] + [p.ptr for p in parts[:4]])
