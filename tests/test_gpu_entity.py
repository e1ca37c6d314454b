"""GPU parity of node and face labs (Cubism::EntityType::Node / Face): the
blocks at the upper end of a plus dimension hold one more element, ghosts
below a block count back from the neighbour's own end
(Block/FieldLabLoader.h:217-219, 236-241), face grids hold 3 * ncomp slabs
(Grid/BlockFieldAssembler.h:153-170).  Every block's lab, from the all-blocks
launch and from the one-block FieldLab, against the oracle bit-exact on the
cells the reference writes, with the written set identical."""
import numpy as np
import pytest

import cubismnova_amd as cnv
import oracle
from tests.helpers import defined_mask, face_bc_lists, random_slab

pytestmark = pytest.mark.gpu

KIND = {0: cnv.Base, 1: cnv.Absorbing, 2: cnv.Dirichlet, 3: cnv.Symmetry}


def mk_bc(t):
    k, d, s, v = t
    return KIND[k](d, s, v) if k in (2, 3) else KIND[k](d, s)


def block_of(b, nb):
    return (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))


def check_entity(entity, nb, n, dtype, seed, bclists, begin, end, tensorial, ncomp=1, slab=0):
    begin = [int(begin)] * 3 if np.isscalar(begin) else list(begin)
    end = [int(end)] * 3 if np.isscalar(end) else list(end)
    g = cnv.Cartesian(nb, n, dtype=dtype, ncomp=ncomp, entity=entity)
    plus = g.plus(slab)
    elems = tuple(n[d] + 1 for d in range(3))  # (N+1)^3 elements per block
    assert g.block_stride * g.dtype.itemsize == -(-int(np.prod(elems)) * g.dtype.itemsize // 32) * 32
    data = random_slab(seed, nb, elems, dtype, block_stride=g.block_stride)
    for c in range(g.nslabs):  # other slabs hold other values
        g.upload(c, data if c == slab else -data)
    for b, lst in bclists.items():
        g.setBC(block_of(b, nb), [mk_bc(t) for t in lst], comp=slab)
    f = oracle.Field(data, nb, n, g.block_stride, bclists, plus=plus)
    st = cnv.Stencil(begin, end, tensorial)
    buf, pitch = g.lab_all(st, comp=slab)
    assert pitch == [n[d] + plus[d] + end[d] - 1 - begin[d] for d in range(3)]
    buf.fill(dtype, np.nan)
    g.lab_all(st, comp=slab, out=buf)
    got = buf.download(dtype).reshape(g.nblocks_total, pitch[2], pitch[1], pitch[0])
    lab = cnv.FieldLab()
    lab.allocate(st, [n[d] + 1 for d in range(3)])
    for b in range(g.nblocks_total):
        bi = block_of(b, nb)
        want = oracle.lab_load(f, bi, begin, end, tensorial)
        m = defined_mask(f, bi, begin, end, tensorial)
        a = want.shape
        _, active = g.lab_geometry(st, slab, bi)
        assert list(a[::-1]) == active
        slot = got[b]
        assert np.isnan(slot[a[0]:]).all() and np.isnan(slot[:, a[1]:]).all() and np.isnan(slot[:, :, a[2]:]).all()
        mine = slot[:a[0], :a[1], :a[2]]
        assert np.array_equal(~np.isnan(mine), m), ("written set differs", bi)
        assert np.array_equal(mine[m], want[m]), ("values differ", bi)
        one = lab.loadData(bi, g, comp=slab)
        assert one.shape == a
        assert np.array_equal(~np.isnan(one), m), ("one-block written set differs", bi)
        assert np.array_equal(one[m], want[m]), ("one-block values differ", bi)
        e = g.block_extent(bi, slab)
        assert lab.getActiveRange()[1] == e
        assert one[-begin[2]:-begin[2] + e[2], -begin[1]:-begin[1] + e[1], -begin[0]:-begin[0] + e[0]].shape == \
            (e[2], e[1], e[0])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tensorial", [False, True])
@pytest.mark.parametrize("sten", [(-1, 2), (-2, 3), ((-1, -2, -3), (3, 2, 1))])
def test_node_lab_periodic(gpu, dtype, tensorial, sten):
    check_entity("node", (3, 2, 2), (5, 4, 6), dtype, 31, {}, sten[0], sten[1], tensorial)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tensorial", [False, True])
def test_node_lab_mixed_bcs(gpu, dtype, tensorial):
    nb, n = (2, 3, 2), (4, 5, 4)
    bcl = face_bc_lists(nb, [(1, 0, 0, 0.0), (3, 0, 1, -1.0), (2, 1, 0, 0.5), (1, 2, 1, 0.0)])
    check_entity("node", nb, n, dtype, 32, bcl, (-2, -2, -2), (3, 3, 3), tensorial)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tensorial", [False, True])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_face_lab_each_direction(gpu, dtype, tensorial, d):
    # two components: slab d*2 + 1 is component 1 of the faces normal to d
    nb, n = (2, 2, 3), (4, 3, 5)
    bcl = face_bc_lists(nb, [(2, (d + 1) % 3, 0, 1.25), (1, d, 1, 0.0)])
    check_entity("face", nb, n, dtype, 40 + d, bcl, (-2, -1, -2), (3, 2, 2), tensorial, ncomp=2, slab=2 * d + 1)


def test_node_one_block_wraps_to_itself(gpu):
    check_entity("node", (1, 1, 1), (4, 4, 4), np.float64, 50, {}, (-2, -2, -2), (3, 3, 3), True)


def test_face_grid_slabs_and_face_bcs(gpu):
    g = cnv.Cartesian((2, 2, 2), (4, 4, 4), entity="face", ncomp=2)
    assert g.nslabs == 6
    assert g.block_stride == 128  # 5^3 doubles rounded up to 32 bytes
    g.addFaceBC(cnv.Dirichlet(0, 0, 1.0))
    for s in range(6):
        assert len(g.getBC((0, 1, 1), comp=s)) == 1
    g.download(5)
    with pytest.raises(cnv.CnvError):
        g.download(6)


@pytest.mark.parametrize("entity", ["node", "face"])
def test_operators_take_cell_grids_only(gpu, entity):
    a = cnv.Cartesian((2, 2, 2), (8, 8, 8), entity=entity)
    b = cnv.Cartesian((2, 2, 2), (8, 8, 8), entity=entity)
    with pytest.raises(cnv.CnvError):
        cnv.laplacian(a, b)
    with pytest.raises(cnv.CnvError):
        cnv.diffusion_step(a, b, 0.1)
