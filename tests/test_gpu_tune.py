"""GPU tests of the autotuner and of explicitly chosen plans.

Every plan the tuner may pick is checked against the oracle first: the
table kernel in each tile configuration the tuner times, and the packed
panels with and without merged result blocks and the persistent schedule
(over 64, 32, 16 and 8 workgroups per XCD), in 128 x 128 and 64 x 64 tiles.
The tuner must leave the caller's blocks untouched, return a plan that
matches the oracle, remember its choice by batch signature (also through a
saved and reloaded table) and never pick a plan slower than the untuned one
beyond timing noise."""
import numpy as np
import pytest

from test_gpu_batch import oracle_batch, per_block_err

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu
TUNER_CONFIGS = [0, 1, 2, 3, 5, 6, 11, 12, 14]
MAKERS = {"ladder": lambda: wl.ladder(),
          "ring-uneven": lambda: wl.ring(o_segs=(9, 6), v_segs=(14, 11, 8))}


def contractions(w):
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    for k in ("D", "L", "R"):
        if w[k].dev is None:
            w[k].to_device()
    return [sg.make_contraction(ptrn, L.addr(lb), L.dims(lb), R.addr(rb), R.dims(rb), D.addr(db),
                                D.dims(db), accumulate=True) for db, lb, rb in w["entries"]]


def prefill(w, seed=5):
    rng = np.random.default_rng(seed)
    w["D"].host[:] = rng.standard_normal(w["D"].host.size)
    w["D"].dev = None


@pytest.mark.parametrize("make", list(MAKERS), ids=list(MAKERS))
@pytest.mark.parametrize("cfg", TUNER_CONFIGS)
def test_table_config_vs_oracle(make, cfg):
    w = MAKERS[make]()
    prefill(w)
    print("tables", make, cfg, flush=True)
    b = sg.plan_choice(contractions(w), "tables", config=cfg)
    assert b.kind == "tables"
    b.execute()
    sg.synchronize()
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w)) <= 1e-13
    b.destroy()


@pytest.mark.parametrize("make", list(MAKERS), ids=list(MAKERS))
@pytest.mark.parametrize("cfg", [0, 11])
@pytest.mark.parametrize("accumulate", [True, False], ids=["acc", "assign"])
def test_table_pair_groups_vs_oracle(make, cfg, accumulate):
    """The table kernel with each result block's pairs cut into groups (tail
    1): the groups' workspace copies are added into the block after the GEMM.
    Accumulating into a prefilled destination and assigning (the first
    entry of each block assigns) both match the oracle."""
    w = MAKERS[make]()
    prefill(w)
    print("tables, pair groups", make, cfg, accumulate, flush=True)
    cs = contractions(w)
    host_d = None
    if not accumulate:
        first = {}
        for i, (db, lb, rb) in enumerate(w["entries"]):
            if db not in first:
                first[db] = i
                cs[i].accumulate = 0
        host_d = np.zeros_like(w["D"].host)
    b = sg.plan_choice(cs, "tables", config=cfg, tail=1)
    assert b.kind == "tables"
    b.execute()
    sg.synchronize()
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w, host_d=host_d)) <= 1e-13
    # a second execution repeats the bits of the first on the same input
    got = w["D"].dev.download()
    w["D"].dev.upload(w["D"].host if accumulate else np.zeros_like(w["D"].host))
    b.execute()
    sg.synchronize()
    assert np.array_equal(w["D"].dev.download(), got)
    b.destroy()


@pytest.mark.parametrize("make", list(MAKERS), ids=list(MAKERS))
@pytest.mark.parametrize("merge,tail,tile", [(1, 1, -1), (1, 0, -1), (0, 1, -1), (0, 0, -1),
                                            (0, 2, -1), (1, 2, -1), (0, 3, -1), (1, 4, -1),
                                            (1, 1, 64), (0, 2, 64), (0, 0, 64), (1, 4, 64)])
def test_packed_choice_vs_oracle(make, merge, tail, tile):
    w = MAKERS[make]()
    prefill(w)
    b = sg.plan_choice(contractions(w), "packed", merge=merge, tail=tail, tile=tile)
    assert b.kind == "packed"
    b.execute()
    sg.synchronize()
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w)) <= 1e-13
    b.destroy()


def test_tuner_picks_a_plan_and_remembers_it(tmp_path):
    sg.tune_clear()
    w = wl.ladder()
    prefill(w)
    cs = contractions(w)
    before = w["D"].dev.download()
    b, res = sg.tune(cs)
    print(res, flush=True)
    assert np.array_equal(w["D"].dev.download(), before)  # tuning ran on scratch blocks
    assert res["candidates"] >= 10
    assert res["default_ms"] > 0 and res["best_ms"] <= 1.10 * res["default_ms"]
    kind = sg.PATHS[res["path"]]
    assert b.kind == kind
    b.execute()
    sg.synchronize()
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w)) <= 1e-13
    b.destroy()
    # the choice is remembered by signature: the same batch shape on other blocks
    w2 = wl.ladder(seed=9)
    prefill(w2, seed=6)
    assert sg.Batch(contractions(w2)).kind == kind
    path = tmp_path / "tune.txt"
    sg.tune_save(path)
    sg.tune_clear()
    sg.tune_load(path)
    b2 = sg.Batch(contractions(w2))
    assert b2.kind == kind
    b2.execute()
    sg.synchronize()
    assert per_block_err(w2, w2["D"].dev.download(), oracle_batch(w2)) <= 1e-13
    b2.destroy()
    sg.tune_clear()


def test_batch_from_a_temporary_workload_keeps_its_blocks():
    """A batch tuned from the contractions of a workload nobody else holds
    keeps that workload (and its device arrays) alive: executing it after a
    garbage collection reads and writes live blocks, and matches the oracle."""
    import gc
    sg.tune_clear()
    b, res = sg.tune(wl.contractions_for(wl.ladder()))
    gc.collect()
    print("temporary workload", res, flush=True)
    b.execute()
    sg.synchronize()
    w = b.contractions.owner
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w)) <= 1e-13
    b.destroy()
    sg.tune_clear()


def test_tuner_split_path():
    sg.tune_clear()
    w = wl.ring(o_segs=(9, 6), v_segs=(14, 11, 8))
    prefill(w)
    b, res = sg.tune(contractions(w), moduli=13)
    # tile -1: the untuned plan was kept (no candidate beat it by 2%)
    assert res["path"] == 2 and b.kind == "split" and res["tile"] in (-1, 128, 256, 257)
    assert res["best_ms"] <= res["default_ms"]
    b.execute()
    sg.synchronize()
    assert per_block_err(w, w["D"].dev.download(), oracle_batch(w)) <= 1e-12
    b.destroy()
    sg.tune_clear()


def test_tune_load_rejects_malformed(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("# header\nzz 1 2\n")
    with pytest.raises(sg.SipGpuError):
        sg.tune_load(p)
    with pytest.raises(sg.SipGpuError):
        sg.tune_load(tmp_path / "missing.txt")


def test_plan_choice_rejects_unknown_tail_and_tile():
    w = wl.ladder()
    cs = contractions(w)
    for tail, tile in ((-2, -1), (5, -1), (1, 128), (1, 257)):
        with pytest.raises(sg.SipGpuError) as e:
            sg.plan_choice(cs, "packed", tail=tail, tile=tile)
        assert e.value.code == -107  # SIPGPU_EARG
