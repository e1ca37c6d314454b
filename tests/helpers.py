"""Shared helpers for the tests: seeded inputs and reference-layout fields."""
import numpy as np

import oracle


def splitmix64(seed, n):
    """Portable 64-bit generator (Steele et al. 2014); returns uint64 array."""
    out = np.empty(n, dtype=np.uint64)
    x = np.uint64(seed)
    with np.errstate(over="ignore"):
        g = np.uint64(0x9E3779B97F4A7C15)
        for i in range(n):
            x = x + g
            z = x
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            out[i] = z ^ (z >> np.uint64(31))
    return out


def uniform(seed, n, lo=-1.0, hi=1.0):
    """53-bit doubles in [lo, hi) from a numpy PCG64 stream (seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return lo + (hi - lo) * rng.random(n)


def iota_field(nblocks, cells, dtype=np.float64):
    nbt = int(np.prod(nblocks))
    n = int(np.prod(cells))
    return np.arange(nbt * n, dtype=np.float64).astype(dtype)


def random_slab(seed, nblocks, cells, dtype=np.float64, block_stride=None):
    nbt = int(np.prod(nblocks))
    n = int(np.prod(cells))
    bs = n if block_stride is None else block_stride
    slab = np.zeros((nbt, bs), dtype=dtype)
    slab[:, :n] = uniform(seed, nbt * n).reshape(nbt, n).astype(dtype)
    return slab.reshape(-1)


def defined_mask(field, bidx, begin, end, tensorial=False, apply_bc=True, extern=None):
    """Cells the reference writes: compare two oracle loads with different poison."""
    a = oracle.lab_load(field, bidx, begin, end, tensorial, apply_bc, extern, poison=1.0e300)
    b = oracle.lab_load(field, bidx, begin, end, tensorial, apply_bc, extern, poison=-7.0e299)
    return a == b


def face_bc_lists(nblocks, face_bcs):
    """Per-block BC lists built the way an application attaches BCs: for each
    (kind, dir, side, value) in order, append it to every block on that face."""
    nbx, nby, nbz = nblocks
    lists = {}
    for (k, d, s, v) in face_bcs:
        for bz in range(nbz):
            for by in range(nby):
                for bx in range(nbx):
                    bi = (bx, by, bz)
                    if bi[d] == (0 if s == 0 else nblocks[d] - 1):
                        lists.setdefault(bx + nbx * (by + nby * bz), []).append((k, d, s, v))
    return lists


def cells_of(slab, nblocks, cells, block_stride):
    """Drop the per-block padding of a slab (block_bytes rounded up to 32)."""
    nbt = int(np.prod(nblocks))
    n = int(np.prod(cells))
    return np.asarray(slab).reshape(nbt, block_stride)[:, :n]


def assert_same(got, want, what=""):
    """Bit-exact comparison with a short message."""
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if not np.array_equal(got, want):
        d = np.abs(got.astype(np.float64) - want.astype(np.float64))
        bad = int(np.count_nonzero(got != want))
        raise AssertionError("%s: %d of %d values differ, max |diff| = %g at flat index %d"
                             % (what, bad, got.size, float(np.nanmax(d)), int(np.nanargmax(d))))
