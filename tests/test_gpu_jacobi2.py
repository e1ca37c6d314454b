"""GPU parity of the two-sweep Jacobi launch (cnv_jacobi_sweep2).

Two sweeps in one launch must equal two single sweeps bit for bit: checked
against two oracle sweeps (periodic, Dirichlet, Symmetry and Absorbing faces
and mixes of them on 3 x 2 x 2 blocks whose spacings differ per direction, a
tanh-stretched z), against two GPU single sweeps for every workgroup size and
z-planes-per-tile of each block edge (tiles split in z, ragged last tile), on
blocks taller than wide, on axes of one block and blocks of 1 to 3 planes, and
on the 512^3 grid of 64^3 blocks.  One cycle of the solver (refined and
all-FP64) equals the same cycle done with single sweeps.  BC lists the launch
does not cover are refused."""
import numpy as np
import pytest

import cubismnova_amd as cnv
import oracle
from cubismnova_amd import stretch
from cubismnova_amd.layout import global_to_blocks
from tests.helpers import assert_same, face_bc_lists

pytestmark = pytest.mark.gpu

# face BC sets as (kind, dir, side, value): 1 Absorbing, 2 Dirichlet, 3 Symmetry (value = sign)
BC2 = {
    "periodic": [],
    "dirichlet": [(2, d, s, 0.25 * (1 + d) * (1 if s else -1)) for d in range(3) for s in (0, 1)],
    "x_periodic": [(2, 1, 0, 0.5), (2, 1, 1, -0.75), (2, 2, 0, 1.0), (2, 2, 1, 0.0)],
    "z_periodic": [(2, 0, 0, 0.0), (2, 0, 1, 2.0), (2, 1, 0, -1.0), (2, 1, 1, 0.0)],
    "symmetric": [(3, 0, 0, 1.0), (3, 0, 1, -1.0), (3, 1, 0, -1.0), (3, 1, 1, -1.0), (3, 2, 0, 1.0),
                  (3, 2, 1, -1.0)],
    "absorbing": [(1, d, s, 0.0) for d in range(3) for s in (0, 1)],
    "mixed_all": [(3, 1, 0, -1.0), (1, 1, 1, 0.0), (2, 2, 0, 0.5), (3, 2, 1, 1.0)],
    "z_mirror": [(2, 0, 0, 0.25), (2, 0, 1, -0.5), (1, 2, 0, 0.0), (3, 2, 1, -1.0)],
}
MAKE = {1: lambda d, s, v: cnv.Absorbing(d, s), 2: lambda d, s, v: cnv.Dirichlet(d, s, v),
        3: lambda d, s, v: cnv.Symmetry(d, s, v)}


def bgrid(nb, n, dtype, bcs, zf):
    g = cnv.Cartesian(nb, n, dtype=dtype)
    for (k, d, s, v) in bcs:
        g.addFaceBC(MAKE[k](d, s, v))
    cnv.set_zfaces(g, zf)
    return g


def clear(g):
    code = 1 if g.dtype == np.float64 else 0
    cnv.call("cnv_fill", g.data_ptr(0), code, 0.0, g.component_bytes // g.dtype.itemsize, None)


def setup(nb, n, dtype, bcs, seed):
    N = [nb[d] * n[d] for d in range(3)]
    zf = stretch.tanh_faces(N[2], 2.0)
    u, f, out, mid = (bgrid(nb, n, dtype, bcs, zf) for _ in range(4))
    rng = np.random.Generator(np.random.PCG64(seed))
    u0 = rng.standard_normal((N[2], N[1], N[0])).astype(dtype)
    f0 = rng.standard_normal((N[2], N[1], N[0])).astype(dtype)
    u.upload_global(0, u0)
    f.upload_global(0, f0)
    return u, f, out, mid, u0, f0, zf


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("bcs", list(BC2), ids=list(BC2))
def test_sweep2_matches_two_oracle_sweeps(gpu, dtype, bcs):
    nb, n = (3, 2, 2), (16, 16, 16)
    u, f, out, mid, u0, f0, zf = setup(nb, n, dtype, BC2[bcs], 11 + len(bcs))
    NX, NY, NZ = 48, 32, 32
    czm, czp, _, _ = stretch.metric(zf)
    h = [1.0 / NX, 1.0 / NY, 1.0 / NZ]
    lists = face_bc_lists(nb, BC2[bcs])
    fs = global_to_blocks(f0, nb, n)
    r1 = oracle.jacobi(oracle.Field(global_to_blocks(u0, nb, n), nb, n, bcs=lists), fs, czm, czp, h, 2.0 / 3.0, 0)
    r2 = oracle.jacobi(oracle.Field(np.asarray(r1), nb, n, bcs=lists), fs, czm, czp, h, 2.0 / 3.0, 0)
    cnv.jacobi_sweep2(u, f, out, 2.0 / 3.0)
    assert_same(out.download(0), r2, "two sweeps in one launch")
    cnv.jacobi_sweep(u, f, mid, 2.0 / 3.0)
    assert_same(mid.download(0), r1, "first single sweep")


def tiles2():
    return [(wg, zc, ys) for wg in (64, 128, 256, 512, 1024) for zc in (8, 16, 24, 32, 64) for ys in (1, 2, 4)]


@pytest.mark.parametrize("bcname", ["x_periodic", "mixed_all"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("edge", [16, 32, 64])
def test_every_sweep2_tile_matches_two_sweeps(gpu, dtype, edge, bcname):
    nb, n = (2, 2, 2), (edge,) * 3
    bcs = BC2[bcname]
    u, f, out, mid, _, _, zf = setup(nb, n, dtype, bcs, 5 * edge)
    ref = bgrid(nb, n, dtype, bcs, zf)
    cnv.jacobi_sweep(u, f, mid, 2.0 / 3.0)
    cnv.jacobi_sweep(mid, f, ref, 2.0 / 3.0)
    want = ref.download(0)
    saved = cnv.get_tune_jacobi2(dtype, edge)
    try:
        ran = 0
        splits = set()
        for wg, zc, ys in tiles2():
            try:
                cnv.set_tune_jacobi2(dtype, edge, wg, zc, ys)
            except cnv.CnvError:
                continue
            clear(out)
            cnv.jacobi_sweep2(u, f, out, 2.0 / 3.0)
            assert_same(out.download(0), want, "tile %d x %d, %d per plane" % (wg, zc, ys))
            ran += 1
            splits.add(ys)
        # the parts of a plane with a tile: a workgroup needs a thread per
        # halo role and at most as many thread rows as the part has rows;
        # blocks of 16^2 are not cut in quarters
        want_splits = {(np.float32, 16): {1}, (np.float64, 16): {1, 2}}.get((dtype, edge), {1, 2, 4})
        assert ran >= 4 and splits == want_splits
    finally:
        cnv.set_tune_jacobi2(dtype, edge, *saved)
    with pytest.raises(cnv.CnvError):
        cnv.set_tune_jacobi2(dtype, edge, 96, 16)
    with pytest.raises(cnv.CnvError):
        cnv.set_tune_jacobi2(dtype, edge, 256, 16, 3)
    if edge == 16:
        with pytest.raises(cnv.CnvError):
            cnv.set_tune_jacobi2(dtype, edge, 64, 16, 4)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sweep2_tall_blocks_ragged_tiles(gpu, dtype):
    """blocks of 16 x 16 x 40 in a 2 x 3 x 2 periodic grid, 16 planes per
    tile: tiles of 16, 16 and 8 planes, z neighbours across block rows"""
    nb, n = (2, 3, 2), (16, 16, 40)
    u, f, out, mid, _, _, zf = setup(nb, n, dtype, [], 99)
    ref = bgrid(nb, n, dtype, [], zf)
    cnv.jacobi_sweep(u, f, mid, 0.8)
    cnv.jacobi_sweep(mid, f, ref, 0.8)
    saved = cnv.get_tune_jacobi2(dtype, 16)
    try:
        cnv.set_tune_jacobi2(dtype, 16, saved[0], 16)
        cnv.jacobi_sweep2(u, f, out, 0.8)
        assert_same(out.download(0), ref.download(0), "tall blocks, 16-plane tiles")
    finally:
        cnv.set_tune_jacobi2(dtype, 16, *saved)
    clear(out)
    cnv.jacobi_sweep2(u, f, out, 0.8)
    assert_same(out.download(0), ref.download(0), "tall blocks, default tiles")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sweep2_full_size_512(gpu, dtype):
    """The 512^3 stretched-z grid of 8 x 8 x 8 blocks of 64^3 with Dirichlet 0
    faces: one two-sweep launch equals two single sweeps in every cell."""
    from cubismnova_amd import workloads
    N, nb, n = 512, (8, 8, 8), (64, 64, 64)
    zf = stretch.tanh_faces(N, 2.0)
    _, _, _, zc = stretch.metric(zf)
    D0 = [(2, d, s, 0.0) for d in range(3) for s in (0, 1)]
    u, f, a, b = (bgrid(nb, n, dtype, D0, zf) for _ in range(4))
    rng = np.random.Generator(np.random.PCG64(2))
    u.upload_global(0, rng.random((N, N, N), dtype=np.float32 if dtype == np.float32 else np.float64).astype(dtype))
    f.upload_global(0, workloads.poisson_source(N, zc).astype(dtype))
    cnv.jacobi_sweep(u, f, a, 2.0 / 3.0)
    cnv.jacobi_sweep(a, f, b, 2.0 / 3.0)
    clear(a)
    cnv.jacobi_sweep2(u, f, a, 2.0 / 3.0)
    d2, r2, mx = cnv.grid_compare(a, b)
    assert d2 == 0.0 and mx == 0.0 and r2 > 0.0


def test_sweep2_refuses_other_bcs(gpu):
    nb, n = (2, 2, 2), (16, 16, 16)
    zf = stretch.uniform_faces(32)

    def g(add):
        x = cnv.Cartesian(nb, n, dtype=np.float32)
        add(x)
        cnv.set_zfaces(x, zf)
        return x

    cases = [
        lambda x: [x.addFaceBC(cnv.Symmetry(0, s, 0.0)) for s in (0, 1)],  # a mirror of sign 0
        lambda x: x.addFaceBC(cnv.Dirichlet(1, 0, 1.0)),  # one Dirichlet face, the other periodic
        lambda x: x.addFaceBC(cnv.Absorbing(2, 1)),        # one Absorbing face, the other periodic
    ]
    for add in cases:
        u, f, out = g(add), g(add), g(add)
        with pytest.raises(cnv.CnvError):
            cnv.jacobi_sweep2(u, f, out, 2.0 / 3.0)
    u = g(lambda x: None)
    with pytest.raises(cnv.CnvError):
        cnv.jacobi_sweep2(u, u, u, 2.0 / 3.0)  # output aliases the input


THIN = [
    ((1, 1, 1), (16, 16, 16), "periodic"),
    ((1, 1, 1), (16, 16, 16), "dirichlet"),
    ((1, 3, 2), (16, 16, 8), "x_periodic"),
    ((2, 2, 1), (16, 16, 2), "periodic"),
    ((2, 2, 1), (16, 16, 2), "dirichlet"),
    ((2, 1, 1), (16, 16, 1), "dirichlet"),
    ((1, 2, 3), (32, 32, 3), "z_periodic"),
    ((1, 1, 1), (16, 16, 16), "symmetric"),
    ((2, 2, 1), (16, 16, 2), "absorbing"),
    ((2, 1, 1), (16, 16, 1), "absorbing"),
    ((1, 3, 2), (16, 16, 8), "mixed_all"),
    ((2, 2, 3), (32, 32, 2), "z_mirror"),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nb,n,bcs", THIN, ids=["%s-%s-%s" % ("x".join(map(str, a)), "x".join(map(str, b)), c)
                                               for a, b, c in THIN])
def test_sweep2_one_block_axes_and_thin_blocks(gpu, dtype, nb, n, bcs):
    """Axes of one block (periodic neighbours are the block itself) and
    blocks of 1, 2 and 3 planes (the two-plane z halo spans several blocks
    or wraps): one two-sweep launch equals two single sweeps, and two oracle
    sweeps on the single-block grids."""
    u, f, out, mid, u0, f0, zf = setup(nb, n, dtype, BC2[bcs], 3 + n[2])
    ref = bgrid(nb, n, dtype, BC2[bcs], zf)
    cnv.jacobi_sweep(u, f, mid, 2.0 / 3.0)
    cnv.jacobi_sweep(mid, f, ref, 2.0 / 3.0)
    cnv.jacobi_sweep2(u, f, out, 2.0 / 3.0)
    assert_same(out.download(0), ref.download(0), "two sweeps in one launch")
    if nb == (1, 1, 1):
        N = [nb[d] * n[d] for d in range(3)]
        czm, czp, _, _ = stretch.metric(zf)
        h = [1.0 / N[0], 1.0 / N[1], 1.0 / N[2]]
        lists = face_bc_lists(nb, BC2[bcs])
        fs = global_to_blocks(f0, nb, n)
        r1 = oracle.jacobi(oracle.Field(global_to_blocks(u0, nb, n), nb, n, bcs=lists), fs, czm, czp, h,
                           2.0 / 3.0, 0)
        r2 = oracle.jacobi(oracle.Field(np.asarray(r1), nb, n, bcs=lists), fs, czm, czp, h, 2.0 / 3.0, 0)
        assert_same(out.download(0), r2, "two oracle sweeps")


def test_sweep2_refuses_single_plane_periodic_z(gpu):
    """a periodic z of one plane in all would wrap the two-plane halo twice"""
    nb, n = (2, 1, 1), (16, 16, 1)
    zf = stretch.uniform_faces(1)
    u, f, out = (bgrid(nb, n, np.float32, [], zf) for _ in range(3))
    with pytest.raises(cnv.CnvError):
        cnv.jacobi_sweep2(u, f, out, 2.0 / 3.0)


@pytest.mark.parametrize("refine", [False, True], ids=["fp64", "refined"])
@pytest.mark.parametrize("S", [2, 4, 6, 10])
@pytest.mark.parametrize("bcs", ["dirichlet", "mixed_all"])
def test_solver_cycle_equals_single_sweeps(gpu, refine, S, bcs):
    """One cycle of cnv_jacobi_solve, whose sweeps run two per launch (and,
    in the all-FP64 solve, two single sweeps when S / 2 is odd), leaves u
    bit-identical to the same cycle done with single sweeps through the
    public entry points."""
    nb, n = (2, 2, 2), (16, 16, 16)
    N = 32
    zf = stretch.tanh_faces(N, 2.0)
    rng = np.random.Generator(np.random.PCG64(S + 7 * refine))
    u0 = rng.standard_normal((N, N, N))
    f0 = rng.standard_normal((N, N, N))
    hom = [(k, d, s, 0.0 if k == 2 else v) for (k, d, s, v) in BC2[bcs]]
    u, f, uref = (bgrid(nb, n, np.float64, BC2[bcs], zf) for _ in range(3))
    u.upload_global(0, u0)
    uref.upload_global(0, u0)
    f.upload_global(0, f0)
    if refine:
        work = cnv.Cartesian(nb, n, dtype=np.float32, ncomp=3)
        cnv.set_zfaces(work, zf)
        # e (comps 0, 1) with the homogeneous BCs, r (comp 2)
        wref = bgrid(nb, n, np.float32, hom, zf)
        e = [bgrid(nb, n, np.float32, hom, zf) for _ in range(2)]
        cnv.jacobi_defect(uref, f, wref)
        clear(e[0])
        for k in range(S):
            cnv.jacobi_sweep(e[k % 2], wref, e[1 - k % 2], 2.0 / 3.0)
        cnv.add_correction(uref, e[S % 2])
    else:
        work = cnv.Cartesian(nb, n, dtype=np.float64, ncomp=2)
        cnv.set_zfaces(work, zf)
        w = bgrid(nb, n, np.float64, BC2[bcs], zf)
        for k in range(S):
            a, b = (uref, w) if k % 2 == 0 else (w, uref)
            cnv.jacobi_sweep(a, f, b, 2.0 / 3.0)
    cyc, sw, rel, _ = cnv.jacobi_solve(u, f, work, refine=refine, sweeps_per_cycle=S, max_cycles=1, tol=0.0)
    assert cyc == 1 and sw == S
    assert_same(u.download(0), uref.download(0), "solver cycle of %d sweeps" % S)
