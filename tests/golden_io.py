"""Access to the reference golden vectors (tests/golden/reference_vectors.npz)."""
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_z = None


def vectors():
    global _z
    if _z is None:
        _z = np.load(os.path.join(HERE, "reference_vectors.npz"), allow_pickle=False)
    return _z


def manifest():
    with open(os.path.join(HERE, "MANIFEST.json")) as f:
        return json.load(f)


def keys(prefix):
    z = vectors()
    return sorted({k.split("_")[0] for k in z.files if k.startswith(prefix)})


def cyclic(n, start):
    """fill_block_cyclic: start, start+1, ..., 20, 1, 2, ... (qm/tests/fill_block_cyclic.f)."""
    k = (np.arange(n) + (int(start) - 1)) % 20
    return (k + 1).astype(np.float64)


def sequential(n, start):
    return start + np.arange(n, dtype=np.float64)


def pattern_case(key):
    z = vectors()
    dr, lr, rr = [int(x) for x in z[key + "_ranks"]]
    aces = [int(x) for x in z[key + "_aces"]][:max(dr, 0) + lr + rr]
    return dr, lr, rr, aces, z[key + "_ptrn"], int(z[key + "_ierr"][0])


def permute_case(key):
    z = vectors()
    dims = [int(x) for x in z[key + "_dims"]]
    transp = [int(x) for x in z[key + "_transp"]]
    n = int(np.prod(dims))
    if key + "_in" in z.files:
        x = z[key + "_in"]
    else:
        kind, start = z[key + "_fill"]
        x = sequential(n, start) if kind == 1.0 else cyclic(n, start)
    return dims, transp, x, z[key + "_out"]


def contract_case(key):
    z = vectors()
    dr, lr, rr = [int(x) for x in z[key + "_ranks"]]
    ptrn = [int(x) for x in z[key + "_ptrn"]][:lr + rr]
    ldims = [int(x) for x in z[key + "_ldims"]][:lr]
    rdims = [int(x) for x in z[key + "_rdims"]][:rr]
    ddims = [int(x) for x in z[key + "_ddims"]][:dr]
    if key + "_cyclic" in z.files:
        s0, s1 = z[key + "_cyclic"]
        L = cyclic(int(np.prod(ldims)) if ldims else 1, s0)
        R = cyclic(int(np.prod(rdims)) if rdims else 1, s1)
    else:
        L, R = z[key + "_l"], z[key + "_r"]
    pre = float(z[key + "_prefill"][0])
    d0 = np.full(int(np.prod(ddims)) if ddims else 1, pre)
    return dict(ptrn=ptrn, ldims=ldims, rdims=rdims, ddims=ddims, l=L, r=R, d0=d0,
                d=z[key + "_d"])
