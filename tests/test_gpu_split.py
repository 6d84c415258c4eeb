"""GPU parity of the split-precision path (FP64 products emulated with INT8
MFMA over pairwise-coprime moduli, recovered by the Chinese remainder theorem).

The oracle is the same as for the FP64 batch: the SIP's own sequence of
tensor_block_contract (assign into a temporary) and block add, in list order.
The split path rounds every operand row (column) to P bits once, so its error
bound is relative to the operands, not to the FP64 rounding:
TOL = 1e-12 relative Frobenius per destination block.
"""
import numpy as np
import pytest

from cases import rel_fro
from test_gpu_batch import oracle_batch, per_block_err

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu
TOL = 1e-12


def run_split(w, moduli, accumulate=True):
    b = wl.batch_for(w, accumulate=accumulate, moduli=moduli)
    b.execute()
    sg.synchronize()
    return b, w["D"].dev.download()


@pytest.mark.parametrize("moduli", [14, 13])
@pytest.mark.parametrize("make", [
    lambda: wl.ladder(),
    lambda: wl.ring(o_segs=(5, 4), v_segs=(10, 10, 7)),
    lambda: wl.ring(o_segs=(15,), v_segs=(30, 15)),
], ids=["ladder", "ring-small", "ring-uneven"])
def test_split_vs_oracle(make, moduli):
    w = make()
    print("split", w["name"], len(w["entries"]), moduli, flush=True)
    rng = np.random.default_rng(3)
    w["D"].host[:] = rng.standard_normal(w["D"].host.size)
    b, got = run_split(w, moduli)
    assert b.moduli == moduli and 30 <= b.bits <= 52
    want = oracle_batch(w)
    assert per_block_err(w, got, want) <= TOL


def test_split_bits_follow_moduli_and_k():
    """P grows with the number of moduli (about 4 bits per modulus) and is
    capped at 52; every choice keeps K * 2^(2P-2) within M / 4."""
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    bits = [wl.batch_for(w, moduli=m).bits for m in (8, 10, 12, 14)]
    assert bits == sorted(bits) and bits[0] < bits[-1]
    assert all(b <= 52 for b in bits)


def test_split_wide_dynamic_range():
    """Elements spanning 2^350 inside one row: each row's scale follows its
    largest element, the terms far below it fall under the FP64 rounding of
    the sum anyway, and the result stays within TOL of the oracle."""
    w = wl.ring(o_segs=(5, 4), v_segs=(10, 7))
    L, R = w["L"], w["R"]
    rng = np.random.default_rng(5)
    L.host[:] *= np.exp2(rng.choice([-200.0, 0.0, 150.0], L.host.size))
    R.host[:] *= 1e-150
    w["D"].host[:] = 0.0
    _, got = run_split(w, 14)
    want = oracle_batch(w)
    assert np.all(np.isfinite(got))
    assert per_block_err(w, got, want) <= TOL


def test_split_row_scaled_exact_relative():
    """Operands scaled by powers of two per destination row keep the result
    within TOL of the oracle block by block (scales are exact)."""
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    w["L"].host[:] *= 2.0 ** 400
    w["R"].host[:] *= 2.0 ** -700
    w["D"].host[:] = 0.0
    _, got = run_split(w, 14)
    want = oracle_batch(w)
    assert per_block_err(w, got, want) <= TOL


def test_split_zero_operands():
    """All-zero operand blocks (no exponent for their rows) give exact zeros,
    and accumulate leaves the destination unchanged."""
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    w["L"].host[:] = 0.0
    w["D"].host[:] = 7.0
    _, got = run_split(w, 14)
    assert np.array_equal(got, w["D"].host)
    _, got = run_split(w, 14, accumulate=False)
    assert np.array_equal(got, np.zeros_like(got))


def test_split_small_integers_exact():
    """Small-integer operands scale exactly, so the emulated product is the
    exact integer sum (representable in FP64 here)."""
    w = wl.ladder()
    rng = np.random.default_rng(7)
    for k in ("L", "R"):
        w[k].host[:] = rng.integers(-64, 65, w[k].host.size).astype(np.float64)
    w["D"].host[:] = 0.0
    _, got = run_split(w, 14)
    want = oracle_batch(w)
    assert np.array_equal(got, want)


def test_split_deterministic_and_linear():
    w = wl.ring(o_segs=(8, 7), v_segs=(16, 16, 9))
    w["D"].host[:] = 0.0
    b = wl.batch_for(w, moduli=14)
    b.execute()
    sg.synchronize()
    first = w["D"].dev.download()
    b.execute()
    sg.synchronize()
    twice = w["D"].dev.download()
    assert np.array_equal(twice.view(np.uint64), (2.0 * first).view(np.uint64)) or \
        rel_fro(twice, 2.0 * first) <= 1e-15
    w["D"].dev.upload(np.zeros_like(first))
    b.execute()
    sg.synchronize()
    again = w["D"].dev.download()
    assert np.array_equal(again.view(np.uint64), first.view(np.uint64))


def test_split_assign_then_accumulate():
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    w["D"].host[:] = 123.0
    for k in ("D", "L", "R"):
        w[k].to_device()
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    cs, seen = [], set()
    for db, lb, rb in w["entries"]:
        first = db not in seen
        seen.add(db)
        cs.append(sg.make_contraction(ptrn, L.addr(lb), L.dims(lb), R.addr(rb), R.dims(rb),
                                      D.addr(db), D.dims(db), accumulate=not first))
    sg.Batch(cs, moduli=14).execute()
    sg.synchronize()
    want = oracle_batch(w, host_d=np.zeros_like(D.host))
    assert per_block_err(w, D.dev.download(), want) <= TOL


def test_split_matches_fp64_batch_on_ring():
    """The benchmark workload: split path against the FP64 MFMA batch (itself
    within 1e-13 of the oracle) on every destination block."""
    w = wl.ring()
    print("split ring full", len(w["entries"]), flush=True)
    b = wl.batch_for(w)
    b.execute()
    sg.synchronize()
    ref = w["D"].dev.download()
    w["D"].dev.upload(np.zeros_like(ref))
    _, got = run_split(w, 14)
    assert per_block_err(w, got, ref) <= TOL


def test_split_rejects_bad_moduli():
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    for m in (0, 5, 15, 16):
        with pytest.raises(sg.SipGpuError) as ei:
            wl.batch_for(w, moduli=m)
        assert ei.value.code == -107


@pytest.mark.parametrize("tile", [None, "128", "257"])
def test_split_long_k_crosses_entry_windows(tile, monkeypatch):
    """One result block summing three pairs of K = 3000 (564 chunks of 16
    bytes of virtual K, more than one staged window of chunk entries), on
    the default GEMM for its K and on the other two pinned."""
    if tile:
        monkeypatch.setenv("SIPGPU_OZ_TILE", tile)
    rng = np.random.default_rng(11)
    M, N, K = 20, 30, 3000
    ptrn, e = sg.get_contraction_ptrn(2, 2, 2, [1, 2, 3, 1, 3, 2])  # D(m,n) = L(k,m) R(k,n)
    assert e == 0
    Ls = [rng.standard_normal(K * M) for _ in range(3)]
    Rs = [rng.standard_normal(K * N) for _ in range(3)]
    dL = [sg.DeviceArray.from_host(x) for x in Ls]
    dR = [sg.DeviceArray.from_host(x) for x in Rs]
    D = sg.DeviceArray(M * N)
    D.zero()
    cs = [sg.make_contraction(ptrn, dL[i].ptr, [K, M], dR[i].ptr, [K, N], D.ptr, [M, N],
                              accumulate=True) for i in range(3)]
    want = sum(Ls[i].reshape(K, M, order="F").T @ Rs[i].reshape(K, N, order="F")
               for i in range(3)).ravel(order="F")
    for moduli in (14, None):
        D.zero()
        b = sg.Batch(cs, moduli=moduli)
        b.execute()
        sg.synchronize()
        assert rel_fro(D.download(), want) <= (1e-12 if moduli else 1e-13)
        b.destroy()


@pytest.mark.parametrize("tile", ["128", "256", "257"])
@pytest.mark.parametrize("make", [
    lambda: wl.ladder(),
    lambda: wl.ring(o_segs=(15,), v_segs=(30, 15)),
], ids=["ladder", "ring-uneven"])
def test_split_both_gemm_tiles(make, tile, monkeypatch):
    """The 128 x 128 and both 256 x 256 INT8 GEMMs (four waves of 128 x 128,
    code 256; eight waves of 64 x 128, code 257) give the same parity;
    SIPGPU_OZ_TILE pins one."""
    monkeypatch.setenv("SIPGPU_OZ_TILE", tile)
    w = make()
    w["D"].host[:] = 0.0
    _, got = run_split(w, 14)
    assert per_block_err(w, got, oracle_batch(w)) <= TOL


@pytest.mark.parametrize("tol", [1e-10, 1e-12, 1e-13])
def test_split_moduli_from_error_target(tol):
    """sipgpu_batch_plan_split_tol takes the fewest moduli with 16 2^-P <=
    tol (one modulus fewer misses it) and the result meets the target
    against the oracle; a target no moduli count reaches is refused."""
    w = wl.ring(o_segs=(15,), v_segs=(30, 15))
    rng = np.random.default_rng(17)
    w["D"].host[:] = rng.standard_normal(w["D"].host.size)
    w["D"].dev = None
    cs = wl.contractions_for(w)
    b = sg.Batch(cs, tol=tol)
    assert b.kind == "split" and 16 * 2.0 ** -b.bits <= tol
    if b.moduli > 6:
        fewer = sg.Batch(cs, moduli=b.moduli - 1)
        assert 16 * 2.0 ** -fewer.bits > tol
        fewer.destroy()
    b.execute()
    sg.synchronize()
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w)) <= tol
    b.destroy()
    with pytest.raises(sg.SipGpuError):
        sg.Batch(cs, tol=1e-17)
