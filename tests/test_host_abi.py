"""The C-ABI library on the CPU: it loads, exports every declared symbol, and
its host-side logic (pattern conversion, index maps, offset tables) matches
the reference.  No device work happens here."""
import ctypes
import os
import re

import numpy as np
import pytest

import golden_io as g
from cases import dims_of, random_case, rel_fro
from checker import Oracle

from aces4_amd import sipgpu as sg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
o = Oracle()


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sipgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?[\s\*]+([A-Za-z_][A-Za-z0-9_]*)"
                       r"\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert "tensor_block_contract__" in names and "sipgpu_batch_plan" in names
    L = sg.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(sg.EXPORTED)


def test_library_is_gfx950_build():
    assert b"gfx950" in sg.lib().sipgpu_version()
    blob = open(sg.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


@pytest.mark.parametrize("key", g.keys("pat"))
def test_pattern_matches_reference(key):
    dr, lr, rr, aces, ptrn, ierr = g.pattern_case(key)
    if dr < 0:
        out = np.zeros(max(lr + rr, 1), np.int32)
        e = ctypes.c_int(0)
        a = np.array(aces, np.int32)
        sg.lib().get_contraction_ptrn_(ctypes.byref(ctypes.c_int(dr)),
                                       ctypes.byref(ctypes.c_int(lr)),
                                       ctypes.byref(ctypes.c_int(rr)),
                                       a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       ctypes.byref(e))
        assert e.value == ierr
        return
    mine, e = sg.get_contraction_ptrn(dr, lr, rr, aces)
    assert e == ierr
    if ierr == 0:
        assert np.array_equal(mine, ptrn)


def test_size_by_shape():
    assert sg.tensor_size_by_shape([3, 4, 5]) == (60, 0)
    assert sg.tensor_size_by_shape([3, 0]) == (0, 2)
    assert sg.tensor_size_by_shape([]) == (1, 0)


def _table_contract(tabs, mnk, L, R):
    """D from the offset tables alone: D[moffD[m]+noffD[n]] = sum_k L[..] R[..]."""
    M, N, K = mnk
    A = L[tabs["koffL"][:K, None] + tabs["moffL"][None, :M]]   # K x M
    B = R[tabs["koffR"][:K, None] + tabs["noffR"][None, :N]]   # K x N
    P = A.T @ B
    D = np.empty(M * N)
    D[(tabs["moffD"][:M, None] + tabs["noffD"][None, :N]).ravel()] = P.ravel()
    return D


def test_index_maps_and_tables_match_oracle():
    rng = np.random.default_rng(21)
    for _ in range(300):
        dl, ll, rl, ext = random_case(rng, allow_scalar=False)
        ptrn, _ = o.get_contraction_ptrn(len(dl), len(ll), len(rl), dl + ll + rl)
        ld, rd, dd = dims_of(ll, ext), dims_of(rl, ext), dims_of(dl, ext)
        mnk, maps, tabs = sg.plan_tables(ptrn, ld, rd, dd)
        ref = o.perms(ptrn, ld, rd, dd)
        # the reference's O2N maps and matrix sizes, bit for bit
        assert (mnk[0], mnk[1], mnk[2]) == (ref["M"], ref["N"], ref["K"])
        assert [int(x) for x in maps["lo2n"]] == ref["lo2n"]
        assert [int(x) for x in maps["ro2n"]] == ref["ro2n"]
        assert [int(x) for x in maps["do2n"]] == ref["do2n"]
        # the tables address every element once and reproduce the contraction
        M, N, K = mnk
        for name, n, size in (("moffD", M, 0), ("noffD", N, 0)):
            assert len(set(tabs[name][:n].tolist())) == n
        dst = (tabs["moffD"][:M, None] + tabs["noffD"][None, :N]).ravel()
        assert sorted(dst.tolist()) == list(range(M * N))
        L = rng.standard_normal(int(np.prod(ld)))
        R = rng.standard_normal(int(np.prod(rd)))
        want, e = o.contract(ptrn, L, ld, R, rd, dd)
        assert e == 0
        assert rel_fro(_table_contract(tabs, mnk, L, R), want) < 1e-14


@pytest.mark.parametrize("key", g.keys("con"))
def test_tables_reproduce_golden_contractions(key):
    c = g.contract_case(key)
    if not (c["ldims"] and c["rdims"] and c["ddims"]):
        pytest.skip("not a block = block * block contraction")
    mnk, maps, tabs = sg.plan_tables(c["ptrn"], c["ldims"], c["rdims"], c["ddims"])
    assert rel_fro(_table_contract(tabs, mnk, c["l"], c["r"]), c["d"]) < 1e-13


def test_pattern_errors_are_reported():
    with pytest.raises(sg.SipGpuError) as ei:
        sg.plan_tables([1, 1], [3], [3], [3])
    assert ei.value.code == -101
    with pytest.raises(sg.SipGpuError) as ei:
        sg.plan_permute([3, 4], [1, 1, 1])
    assert ei.value.code == -103


def test_permute_plans():
    # identity and extent-1 dims collapse to a plain copy
    assert sg.plan_permute([4, 5, 6], [1, 1, 2, 3])[3] == 1
    assert sg.plan_permute([1, 7, 1], [1, 3, 1, 2])[3] == 1
    # (3,4,1,2) on a ring block: two canonical dims, transpose mode
    rank, nU, row, ident = sg.plan_permute([30, 15, 30, 15], [1, 3, 4, 1, 2])
    assert (rank, row, ident) == (2, 0, 0)
    # input-fastest dim stays fastest: row mode
    assert sg.plan_permute([8, 3, 5], [1, 1, 3, 2])[2] == 1


def test_device_count_without_gpu_is_not_an_error():
    assert sg.device_count() >= 0


def test_product_path_fails_loudly_without_gpu():
    if sg.device_count() > 0:
        pytest.skip("a GPU is present")
    _, ierr = sg.tensor_block_contract([1, -1, -2, 2], np.ones(6), [2, 3], np.ones(12), [3, 4],
                                       [2, 4])
    assert ierr != 0
    _, ierr = sg.tensor_block_copy([2, 3], [1, 2, 1], np.ones(6))
    assert ierr != 0


def _slice_offsets(te, se, eb):
    """Tensor offsets of the slice's elements, in slice order (numpy closed form)."""
    idx = np.arange(int(np.prod(te)), dtype=np.int64).reshape(te, order="F")
    return idx[tuple(slice(b, b + n) for b, n in zip(eb, se))].ravel(order="F")


@pytest.mark.parametrize("seed", range(10))
def test_slice_plans_reproduce_offsets(seed):
    rng = np.random.default_rng(seed)
    rank = 1 + seed % 6
    te = [int(v) for v in rng.integers(1, 7, size=rank)]
    se = [int(rng.integers(1, t + 1)) for t in te]
    if seed % 3 == 0:  # full leading dims: they merge
        se[:2] = te[:2]
    eb = [int(rng.integers(0, t - n + 1)) for t, n in zip(te, se)]
    ext, st, base, count, span, tsize = sg.plan_slice(te, se, eb)
    want = _slice_offsets(te, se, eb)
    assert count == want.size and tsize == int(np.prod(te))
    # canonical digits reproduce every offset
    l = np.arange(count, dtype=np.int64)
    off = np.full(count, base, dtype=np.int64)
    for e, s in zip(ext, st):
        off += (l % e) * s
        l //= e
    np.testing.assert_array_equal(off, want)
    assert base == want.min() and span == want.max() - want.min() + 1
    assert all(e > 1 for e in ext)
    assert all(st[i + 1] != st[i] * ext[i] for i in range(len(ext) - 1))  # nothing left to merge


def test_slice_plan_canonical_forms_and_errors():
    assert sg.plan_slice([4, 5, 6], [4, 5, 6], [0, 0, 0])[:2] == ([120], [1])
    assert sg.plan_slice([4, 5, 6], [4, 2, 1], [0, 3, 5])[:3] == ([8], [1], 12 + 100)
    assert sg.plan_slice([4, 5, 6], [1, 1, 1], [3, 4, 5])[:5] == ([], [], 119, 1, 1)
    assert sg.plan_slice([], [], [])[2:] == (0, 1, 1, 1)       # rank 0: one element
    assert sg.plan_slice([4, 5], [0, 5], [0, 0])[3] == 0       # empty slice
    for te, se, eb in (([4, 5], [3, 5], [2, 0]), ([4, 5], [1, 1], [-1, 0]),
                       ([4, 5], [-1, 1], [0, 0])):
        with pytest.raises(sg.SipGpuError) as ei:
            sg.plan_slice(te, se, eb)
        assert ei.value.code == -107
    with pytest.raises(sg.SipGpuError) as ei:
        sg.plan_slice([1 << 16, 1 << 16], [1, 1], [0, 0])
    assert ei.value.code == -104


def test_fastdiv_is_exact(tmp_path):
    """fastdiv.h against the hardware divide: every n < 2^16 for divisors up
    to 300, and boundary and random numerators up to 2^31 - 1 for divisors
    across the 32-bit range."""
    src = tmp_path / "fd.cpp"
    src.write_text(r'''
#include "fastdiv.h"
#include <cstdio>
#include <cstdint>
using namespace sipgpu;
static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint32_t rnd() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return uint32_t(st); }
static int check(uint32_t d, uint32_t n) {
    const FastDiv f = make_fastdiv(d);
    if (fastdiv(n, f) != n / d) { std::printf("d=%u n=%u\n", d, n); return 1; }
    return 0;
}
int main() {
    int bad = 0;
    for (uint32_t d = 1; d <= 300; ++d)
        for (uint32_t n = 0; n < (1u << 16); ++n) bad += check(d, n);
    const uint32_t top = (1u << 31) - 1;
    for (int i = 0; i < 200000; ++i) {
        uint32_t d = 1 + rnd() % top;
        if (i & 1) d = 1 + d % 4096;
        const uint32_t q = top / d;
        bad += check(d, top) + check(d, rnd() % (top + 1u)) + check(d, d - 1) + check(d, d);
        bad += check(d, q * d) + check(d, q * d - 1);
    }
    for (int s = 0; s < 31; ++s) {
        const uint32_t p = 1u << s;
        bad += check(p, top) + check(p + 1, top) + check(p - (s > 0), top);
    }
    std::printf("%d\n", bad);
    return bad != 0;
}
''')
    exe = tmp_path / "fd"
    import subprocess
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "aces4_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]


def test_block_operations_fail_loudly_without_gpu():
    if sg.device_count() > 0:
        pytest.skip("a GPU is present")
    assert sg.tensor_block_init([2, 3], np.zeros(6), 1.0)[1] != 0
    assert sg.tensor_block_slice(np.arange(6.0), [2, 3], [1, 2], [1, 0])[1] != 0
    # argument errors come first and keep the reference's codes
    assert sg.tensor_block_init([2], np.zeros(2), 1.0, rank=-1)[1] == -1
    assert sg.tensor_block_slice(np.arange(6.0), [2, 3], [1, 2], [1, 0], rank=-1)[1] == 1
    assert sg.tensor_block_slice(np.arange(6.0), [2, 3], [0, 2], [1, 0])[1] == -1
    assert sg.tensor_block_slice(np.arange(6.0), [2, 3], [2, 2], [1, 0])[1] == -107


# Every bind(c) name of the reference's tensor module (tensor_dil_omp.F90,
# prototypes tensor_ops_c_prototypes.h:41-178): the library linked in its
# place must export all of them.
REFERENCE_MODULE_SYMBOLS = [
    "tensor_size_by_shape_", "get_contraction_ptrn_", "tensor_block_init__",
    "tensor_block_scale__", "tensor_block_norm2__", "tensor_block_slice__",
    "tensor_block_insert__", "tensor_block_add__", "tensor_block_copy__",
    "tensor_block_contract__",
]


def test_library_exports_the_whole_reference_module():
    L = sg.lib()
    missing = [n for n in REFERENCE_MODULE_SYMBOLS if not hasattr(L, n)]
    assert not missing, missing
    names = declared_functions()
    assert all(n in names for n in REFERENCE_MODULE_SYMBOLS)


def test_native_block_operations_fail_loudly_without_gpu():
    """No CPU fallback: without a device every native block operation returns
    an error (the fake device address is never touched)."""
    if sg.device_count() > 0:
        pytest.skip("a GPU is present")
    fake = 0x100000
    calls = [
        lambda: sg.block_init(fake, [4, 4], 1.0),
        lambda: sg.block_scale(fake, [4, 4], 2.0),
        lambda: sg.block_add(fake, fake + 4096, [4, 4], 0.5),
        lambda: sg.block_norm2(fake, [4, 4], fake + 4096),
        lambda: sg.block_slice(fake, [4, 4], fake + 4096, [2, 2], [1, 1]),
        lambda: sg.block_insert(fake, [4, 4], fake + 4096, [2, 2], [1, 1]),
        lambda: sg.block_add_many(fake, [fake + 4096], [4, 4]),
    ]
    for call in calls:
        with pytest.raises(sg.SipGpuError):
            call()
    # argument errors are reported before the device is looked for
    with pytest.raises(sg.SipGpuError) as ei:
        sg.block_slice(fake, [4, 4], fake + 4096, [2, 2], [3, 0])
    assert ei.value.code == -107


def test_stream_wait_event_rejects_a_null_event():
    rc = sg.lib().sipgpu_stream_wait_event(None, None)
    assert rc == -107


def test_slice_plans_beyond_the_sip_rank():
    """The reference module accepts any rank; canonical forms of up to eight
    dims run, a slice that needs nine is refused with SIPGPU_ERANK."""
    te = [3, 1, 4, 1, 5, 2, 1, 3, 2, 2]
    se = [3, 1, 2, 1, 5, 2, 1, 1, 2, 1]
    eb = [0, 0, 1, 0, 0, 0, 0, 2, 0, 1]
    ext, st, base, count, span, tsize = sg.plan_slice(te, se, eb)
    want = _slice_offsets(te, se, eb)
    l = np.arange(count, dtype=np.int64)
    off = np.full(count, base, dtype=np.int64)
    for e, s in zip(ext, st):
        off += (l % e) * s
        l //= e
    np.testing.assert_array_equal(off, want)
    # nine dims of extent 2 in a 3^9 block, none mergeable: nine canonical dims
    with pytest.raises(sg.SipGpuError) as ei:
        sg.plan_slice([3] * 9, [2] * 9, [0] * 9)
    assert ei.value.code == -102


def test_tune_load_validates_fields(tmp_path):
    """A tuning table line is 'signature path config merge tail tile'; out of
    range fields (tail 5, a split tile on the packed path, an unknown tile)
    are refused, well-formed lines load (host-only, no device work)."""
    good = tmp_path / "good.txt"
    good.write_text("# sipgpu tuning table\n1a 1 -1 0 2 64\n2b 2 -1 1 -1 257\n3c 0 5 -1 -1 -1\n")
    sg.tune_load(good)
    for line in ("4d 1 -1 0 5 -1", "5e 1 -1 0 1 257", "6f 2 -1 0 -1 64", "70 1 -1 2 1 -1"):
        bad = tmp_path / "bad.txt"
        bad.write_text(line + "\n")
        with pytest.raises(sg.SipGpuError):
            sg.tune_load(bad)
    sg.tune_clear()


def test_triples_launch_size_matches_the_header():
    """sipgpu_triples_run cuts a slot range into launches of
    SIPGPU_TRIPLES_LAUNCH_BOXES slots; the Python mirror agrees, and a launch
    of 512-thread workgroups stays below 2^32 work-items."""
    from aces4_amd import triples as tp
    text = open(os.path.join(ROOT, "include", "sipgpu.h")).read()
    m = re.search(r"#define SIPGPU_TRIPLES_LAUNCH_BOXES \(1LL << (\d+)\)", text)
    assert m and tp.LAUNCH_BOXES == 1 << int(m.group(1))
    assert (tp.LAUNCH_BOXES + 7) // 8 * 8 * 512 < 1 << 32
