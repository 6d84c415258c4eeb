"""GPU parity of planned batches (one grouped FP64-MFMA launch per batch).

The oracle of a batch is the SIP's own sequence: for every block pair,
TMP = L * R (assign, tensor_block_contract) then D += TMP (block add), in
list order.  Tolerance 1e-13 relative Frobenius per destination block.
"""
import numpy as np
import pytest

from cases import rel_fro
from checker import Oracle

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu
o = Oracle()
TOL = 1e-13


def oracle_batch(w, entries=None, host_d=None):
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    out = (D.host if host_d is None else host_d).copy()
    for db, lb, rb in (w["entries"] if entries is None else entries):
        tmp, e = o.contract(ptrn, L.block(lb), L.dims(lb), R.block(rb), R.dims(rb), D.dims(db))
        assert e == 0
        n = tmp.size
        out[D.offset[db]:D.offset[db] + n] += tmp
    return out


def per_block_err(w, got, want):
    D = w["D"]
    worst = 0.0
    for db in D.block_ids():
        worst = max(worst, rel_fro(D.block(db, got), D.block(db, want)))
    return worst


@pytest.mark.parametrize("make", [
    lambda: wl.ladder(),
    lambda: wl.ring(o_segs=(5, 4), v_segs=(10, 10, 7)),
    lambda: wl.ring(o_segs=(15,), v_segs=(30, 15)),
], ids=["ladder", "ring-small", "ring-uneven"])
def test_batch_vs_oracle(make):
    w = make()
    print("batch", w["name"], len(w["entries"]), flush=True)
    rng = np.random.default_rng(1)
    w["D"].host[:] = rng.standard_normal(w["D"].host.size)
    b = wl.batch_for(w)
    b.execute()
    sg.synchronize()
    got = w["D"].dev.download()
    want = oracle_batch(w)
    assert per_block_err(w, got, want) <= TOL
    assert abs(b.flops - wl.flops(w)) < 1e-6 * wl.flops(w)


def test_batch_assign_then_accumulate():
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    w["D"].host[:] = 123.0
    for k in ("D", "L", "R"):
        w[k].to_device()
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    cs = []
    seen = set()
    for db, lb, rb in w["entries"]:
        first = db not in seen
        seen.add(db)
        cs.append(sg.make_contraction(ptrn, L.addr(lb), L.dims(lb), R.addr(rb), R.dims(rb),
                                      D.addr(db), D.dims(db), accumulate=not first))
    sg.Batch(cs).execute()
    sg.synchronize()
    want = oracle_batch(w, host_d=np.zeros_like(D.host))
    assert per_block_err(w, D.dev.download(), want) <= TOL


def test_batch_same_alone_and_in_mix():
    """A destination block's result is bitwise the same whether its task runs
    alone or inside a larger batch: tasks never share accumulators."""
    w = wl.ring(o_segs=(8, 7), v_segs=(16, 16, 9))
    for k in ("D", "L", "R"):
        w[k].to_device()
    D = w["D"]
    wl.batch_for(w).execute()
    sg.synchronize()
    mixed = D.dev.download()
    target = next(iter(D.block_ids()))
    alone_entries = [e for e in w["entries"] if e[0] == target]
    D.dev.upload(np.zeros_like(D.host))
    w2 = dict(w, entries=alone_entries)
    wl.batch_for(w2).execute()
    sg.synchronize()
    alone = D.dev.download()
    assert np.array_equal(D.block(target, alone).view(np.uint64),
                          D.block(target, mixed).view(np.uint64))


def test_batch_hazard_rejected():
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    for k in ("D", "L", "R"):
        w[k].to_device()
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    db, lb, rb = w["entries"][0]
    # destination aliases the right operand
    c = sg.make_contraction(ptrn, L.addr(lb), L.dims(lb), R.addr(rb), R.dims(rb), R.addr(rb),
                            D.dims(db))
    with pytest.raises(sg.SipGpuError) as ei:
        sg.Batch([c])
    assert ei.value.code == -108


def test_full_ring_config_properties():
    """The benchmark workload (o=30, v=150, uneven segments): a sample of
    destination blocks against the oracle, and linearity in T."""
    w = wl.ring()
    print("batch ring full", len(w["entries"]), flush=True)
    b = wl.batch_for(w)
    b.execute()
    sg.synchronize()
    got = w["D"].dev.download()
    D = w["D"]
    blocks = list(D.block_ids())
    sample = [blocks[0], blocks[len(blocks) // 2], blocks[-1]]
    for db in sample:
        ents = [e for e in w["entries"] if e[0] == db]
        want = oracle_batch(w, ents)
        assert rel_fro(D.block(db, got), D.block(db, want)) <= TOL
    # linearity: D(T) computed twice (accumulate) equals 2 * D(T)
    b.execute()
    sg.synchronize()
    twice = w["D"].dev.download()
    assert rel_fro(twice, 2.0 * got) <= 1e-15


@pytest.mark.parametrize("make", [
    lambda: wl.ladder(),
    lambda: wl.ring(o_segs=(5, 4), v_segs=(10, 10, 7)),
    lambda: wl.ring(o_segs=(15,), v_segs=(30, 15)),
], ids=["ladder", "ring-small", "ring-uneven"])
def test_packed_path_vs_oracle(make, monkeypatch):
    """The packed FP64 path (operands packed into panels, GEMM over virtual
    K) against the oracle, forced on batches the planner would run on the
    table kernel."""
    monkeypatch.setenv("SIPGPU_FP64_PATH", "packed")
    w = make()
    rng = np.random.default_rng(1)
    w["D"].host[:] = rng.standard_normal(w["D"].host.size)
    b = wl.batch_for(w)
    assert b.kind == "packed"
    b.execute()
    sg.synchronize()
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w)) <= TOL


def test_packed_path_assign_and_long_k(monkeypatch):
    """Assign then accumulate on the packed path, and one block summing three
    pairs of K = 3000 (4500 chunks of virtual K: several windows of chunk
    entries)."""
    monkeypatch.setenv("SIPGPU_FP64_PATH", "packed")
    rng = np.random.default_rng(12)
    M, N, K = 20, 30, 3000
    ptrn, e = sg.get_contraction_ptrn(2, 2, 2, [1, 2, 3, 1, 3, 2])
    Ls = [rng.standard_normal(K * M) for _ in range(3)]
    Rs = [rng.standard_normal(K * N) for _ in range(3)]
    dL = [sg.DeviceArray.from_host(x) for x in Ls]
    dR = [sg.DeviceArray.from_host(x) for x in Rs]
    D = sg.DeviceArray.from_host(np.full(M * N, 7.0))
    cs = [sg.make_contraction(ptrn, dL[i].ptr, [K, M], dR[i].ptr, [K, N], D.ptr, [M, N],
                              accumulate=(i > 0)) for i in range(3)]
    b = sg.Batch(cs)
    assert b.kind == "packed"
    b.execute()
    sg.synchronize()
    want = sum(Ls[i].reshape(K, M, order="F").T @ Rs[i].reshape(K, N, order="F")
               for i in range(3)).ravel(order="F")
    assert rel_fro(D.download(), want) <= TOL


def test_full_ring_runs_packed():
    w = wl.ring()
    b = wl.batch_for(w)
    assert b.kind == "packed"


@pytest.mark.parametrize("path", ["tables", "packed", "split"])
def test_stages_equal_execute(path, monkeypatch):
    """PREPARE, PRODUCT, FINISH run one at a time give bit for bit what one
    sipgpu_batch_execute gives, on each of the three paths; a repeated
    PRODUCT after one PREPARE repeats the same product."""
    if path != "split":
        monkeypatch.setenv("SIPGPU_FP64_PATH", path)
    w = wl.ring(o_segs=(9, 6), v_segs=(14, 11, 8))
    rng = np.random.default_rng(3)
    w["D"].host[:] = rng.standard_normal(w["D"].host.size)
    w["D"].to_device()
    b = wl.batch_for(w, moduli=13 if path == "split" else None)
    assert b.kind == path
    b.execute()
    sg.synchronize()
    once = w["D"].dev.download()
    w["D"].dev.upload(w["D"].host)
    for st in (sg.STAGE_PREPARE, sg.STAGE_PRODUCT, sg.STAGE_FINISH):
        b.execute_stages(st)
    sg.synchronize()
    assert np.array_equal(w["D"].dev.download(), once)
    with pytest.raises(sg.SipGpuError):
        b.execute_stages(8)
