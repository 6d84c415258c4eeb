"""GPU edge cases of planned batches on all three paths (table kernel,
packed FP64 panels, split precision): empty batches, extent-1 indices and
segments, K = 1, row / column vector results, ragged segment lists, and the
size limits the planners refuse.  Results against the oracle (the SIP's own
contract + add sequence) within each path's tolerance."""
import numpy as np
import pytest

from cases import rel_fro
from test_gpu_batch import oracle_batch, per_block_err

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu
PATHS = {"tables": 1e-13, "packed": 1e-13, "split": 1e-12}


def plan(cs, path):
    if path == "split":
        return sg.Batch(cs, moduli=14)
    return sg.plan_choice(cs, path)


@pytest.mark.parametrize("path", list(PATHS))
def test_empty_batch(path):
    b = plan([], path)
    assert b.ntasks == 0 and b.npairs == 0 and b.flops == 0.0
    b.execute()
    for st in (sg.STAGE_PREPARE, sg.STAGE_PRODUCT, sg.STAGE_FINISH):
        b.execute_stages(st)
    sg.synchronize()
    b.destroy()


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("mkn", [(1, 1, 1), (1, 37, 23), (41, 1, 19), (29, 31, 1), (1, 1, 300),
                                 (200, 150, 1)], ids=lambda x: "x".join(map(str, x)))
def test_degenerate_extents(path, mkn):
    """D(m,n) (+)= sum_k L(k,m) R(k,n) with M, N or K equal to 1, assign and
    accumulate entries into the same block."""
    M, N, K = mkn
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    ptrn, e = sg.get_contraction_ptrn(2, 2, 2, [1, 2, 3, 1, 3, 2])
    assert e == 0
    Ls = [rng.standard_normal(K * M) for _ in range(2)]
    Rs = [rng.standard_normal(K * N) for _ in range(2)]
    dL = [sg.DeviceArray.from_host(x) for x in Ls]
    dR = [sg.DeviceArray.from_host(x) for x in Rs]
    D0 = rng.standard_normal(M * N)
    D = sg.DeviceArray.from_host(D0)
    cs = [sg.make_contraction(ptrn, dL[i].ptr, [K, M], dR[i].ptr, [K, N], D.ptr, [M, N],
                              accumulate=True) for i in range(2)]
    b = plan(cs, path)
    b.execute()
    sg.synchronize()
    want = D0 + sum(Ls[i].reshape(K, M, order="F").T @ Rs[i].reshape(K, N, order="F")
                    for i in range(2)).ravel(order="F")
    assert rel_fro(D.download(), want) <= PATHS[path]
    b.destroy()


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("segs", [((5, 1, 7), (17, 3, 29, 1)), ((1,), (2, 1)), ((3, 4), (1, 1, 1))],
                         ids=["ragged", "tiny", "unit-virtuals"])
def test_ragged_segments(path, segs):
    """The ring term over ragged segment lists that include extent-1
    segments: blocks with unit indices in every position."""
    o, v = segs
    w = wl.ring(o_segs=o, v_segs=v, seed=4)
    rng = np.random.default_rng(9)
    w["D"].host[:] = rng.standard_normal(w["D"].host.size)
    w["D"].dev = None
    b = plan(wl.contractions_for(w), path)
    b.execute()
    sg.synchronize()
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w)) <= PATHS[path]
    b.destroy()


@pytest.mark.parametrize("path", list(PATHS))
def test_oversized_result_block_refused(path):
    """A result block of 2^31 elements or more is refused at planning (the
    blocks are never touched: the plan fails before any launch)."""
    x = sg.DeviceArray(64)
    y = sg.DeviceArray(64)
    z = sg.DeviceArray(64)
    ptrn, e = sg.get_contraction_ptrn(2, 2, 2, [1, 2, 3, 1, 3, 2])
    assert e == 0
    c = sg.make_contraction(ptrn, x.ptr, [4, 50000], y.ptr, [4, 50000], z.ptr, [50000, 50000])
    with pytest.raises(sg.SipGpuError):
        plan([c], path)
