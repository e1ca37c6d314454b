"""Loader of the committed known-answer fixtures (tests/golden/lab_fixtures.npz)."""
import json
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lab_fixtures.npz")
KIND = {"base": 0, "absorbing": 1, "dirichlet": 2, "symmetry": 3}


def load():
    z = np.load(PATH, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


def bc_lists(entry):
    """{block id: [(kind int, dir, side, value), ...]} in list order."""
    return {int(b): [(KIND[k], d, s, v) for (k, d, s, v) in lst] for b, lst in entry["bc_lists"].items()}


def elements(entry):
    """Elements per block of a fixture grid: (cells + plus) per dimension."""
    return tuple(c + p for c, p in zip(entry["cells"], plus(entry)))


def plus(entry):
    return tuple(entry.get("plus", (0, 0, 0)))


def extent(entry, bi):
    """Extent of block bi: one more in a plus dimension at the upper end."""
    p, nb = plus(entry), entry["nblocks"]
    return tuple(entry["cells"][d] + (1 if p[d] and bi[d] == nb[d] - 1 else 0) for d in range(3))


def padded(values, nblocks, cells, block_stride, dtype):
    """Block-major values (cells = elements per block) -> slab with the
    grid's block stride."""
    nbt = int(np.prod(nblocks))
    n = int(np.prod(cells))
    out = np.zeros((nbt, block_stride), dtype=dtype)
    out[:, :n] = np.asarray(values).reshape(nbt, n).astype(dtype)
    return out.reshape(-1)


def block_index(b, nb):
    return (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
