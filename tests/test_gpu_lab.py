"""GPU parity of the lab assembly: every block's lab from the HIP kernel
(through the C ABI) against the oracle, bit-exact on the cells the reference
writes, and the set of written cells identical."""
import itertools

import numpy as np
import pytest

import cubismnova_amd as cnv
import oracle
from tests.helpers import defined_mask, face_bc_lists, random_slab

pytestmark = pytest.mark.gpu

KIND = {0: cnv.Base, 1: cnv.Absorbing, 2: cnv.Dirichlet, 3: cnv.Symmetry}


def mk_bc(t):
    k, d, s, v = t
    if k in (2, 3):
        return KIND[k](d, s, v)
    return KIND[k](d, s)


def make_grid(nb, n, dtype, slab, bclists):
    g = cnv.Cartesian(nb, n, dtype=dtype)
    assert g.block_stride == int(np.prod(n))
    g.upload(0, slab)
    for b, lst in bclists.items():
        bi = (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
        g.setBC(bi, [mk_bc(t) for t in lst])
    return g


def check_all_blocks(nb, n, dtype, seed, bclists, begin, end, tensorial):
    g = cnv.Cartesian(nb, n, dtype=dtype)
    slab = random_slab(seed, nb, n, dtype, block_stride=g.block_stride)
    g.upload(0, slab)
    for b, lst in bclists.items():
        bi = (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
        g.setBC(bi, [mk_bc(t) for t in lst])
    f = oracle.Field(slab, nb, n, g.block_stride, bclists)
    st = cnv.Stencil(begin, end, tensorial)
    buf, ext = g.lab_all(st)
    buf.fill(dtype, np.nan)
    g.lab_all(st, out=buf)
    got = buf.download(dtype).reshape(g.nblocks_total, ext[2], ext[1], ext[0])
    for b in range(g.nblocks_total):
        bi = (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
        want = oracle.lab_load(f, bi, begin, end, tensorial)
        m = defined_mask(f, bi, begin, end, tensorial)
        written = ~np.isnan(got[b])
        assert np.array_equal(written, m), ("written set differs", bi)
        assert np.array_equal(got[b][m], want[m]), ("values differ", bi)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tensorial", [False, True])
@pytest.mark.parametrize("sten", [(-1, 2), (-3, 4), (-2, 3), ((-1, -2, -3), (3, 2, 1))])
def test_lab_periodic(gpu, dtype, tensorial, sten):
    check_all_blocks((3, 2, 4), (5, 4, 6), dtype, 21, {}, sten[0], sten[1], tensorial)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tensorial", [False, True])
@pytest.mark.parametrize("g", [1, 3])
def test_lab_mixed_bcs(gpu, dtype, tensorial, g):
    # x periodic, y absorbing, z Dirichlet -0.25 / +0.25
    nb, n = (2, 2, 2), (8, 8, 8)
    bcl = face_bc_lists(nb, [(1, 1, 0, 0.0), (1, 1, 1, 0.0), (2, 2, 0, -0.25), (2, 2, 1, 0.25)])
    check_all_blocks(nb, n, dtype, 7, bcl, -g, g + 1, tensorial)


@pytest.mark.parametrize("tensorial", [False, True])
def test_lab_one_block_both_sides(gpu, tensorial):
    bcl = {0: [(2, 0, 0, -1.0), (2, 0, 1, 2.0), (3, 1, 0, -1.0), (1, 1, 1, 0.0), (3, 2, 1, 0.5)]}
    check_all_blocks((1, 1, 1), (6, 6, 6), np.float64, 8, bcl, -2, 3, tensorial)


def test_lab_order_dependent_corners(gpu):
    for order in itertools.permutations([(2, 0, 0, 5.0), (1, 1, 0, 0.0), (3, 2, 1, -1.0)]):
        bcl = {0: list(order), 1: list(order[::-1])}
        check_all_blocks((2, 1, 1), (4, 4, 4), np.float64, 9, bcl, -1, 2, True)


def test_lab_random_fuzz(gpu):
    rng = np.random.Generator(np.random.PCG64(1234))
    for trial in range(12):
        nb = tuple(int(v) for v in rng.integers(1, 4, size=3))
        n = tuple(int(v) for v in rng.integers(3, 7, size=3))
        dtype = np.float64 if trial % 2 == 0 else np.float32
        nbt = nb[0] * nb[1] * nb[2]
        bcl = {}
        for b in range(nbt):
            lst = []
            for _ in range(int(rng.integers(0, 5))):
                lst.append((int(rng.integers(0, 4)), int(rng.integers(0, 3)), int(rng.integers(0, 2)),
                            float(rng.choice([-1.0, 1.0, 0.5, 2.0]))))
            bcl[b] = lst
        gw = int(rng.integers(1, 3))
        begin = [-int(rng.integers(0, gw + 1)) for _ in range(3)]
        end = [int(rng.integers(1, gw + 2)) for _ in range(3)]
        tens = bool(rng.integers(0, 2))
        check_all_blocks(nb, n, dtype, 100 + trial, bcl, begin, end, tens)


def test_lab_single_block_extern_and_apply(gpu):
    nb, n = (2, 2, 2), (6, 6, 6)
    slab = random_slab(10, nb, n)
    bcl = face_bc_lists(nb, [(2, 0, 0, 3.0), (1, 2, 1, 0.0)])
    g = make_grid(nb, n, np.float64, slab, bcl)
    f = oracle.Field(slab, nb, n, bcs=bcl)
    ext = [(3, 1, 0, -1.0), (2, 0, 1, 7.0)]
    for bi in itertools.product(range(2), repeat=3):
        for apply_bc in (True, False):
            for extern in (None, ext):
                lab = cnv.FieldLab()
                lab.allocate(cnv.Stencil(-2, 3, False), n)
                got = lab.loadData(bi, g, apply_bc=apply_bc,
                                   extern_bc=None if extern is None else [mk_bc(t) for t in extern])
                want = oracle.lab_load(f, bi, -2, 3, False, apply_bc, extern)
                m = defined_mask(f, bi, -2, 3, False, apply_bc, extern)
                assert np.array_equal(~np.isnan(got), m)
                assert np.array_equal(got[m], want[m])


def test_lab_errors(gpu):
    lab = cnv.FieldLab()
    g = cnv.Cartesian(1, 4)
    with pytest.raises(RuntimeError, match="not allocated first"):
        lab.loadData((0, 0, 0), g)
    with pytest.raises(RuntimeError, match="no field loaded"):
        lab.getActiveField()
    with pytest.raises(RuntimeError, match="begin_ must be <= 0"):
        cnv.Stencil(1, 2)
    lab.allocate(cnv.Stencil(-5, 6), [4, 4, 4])
    with pytest.raises(cnv.CnvError):
        lab.loadData((0, 0, 0), g)  # ghost width larger than the block
