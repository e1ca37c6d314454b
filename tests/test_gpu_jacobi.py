"""GPU parity of the stretched-z Jacobi kernels and the refined solver.

Sweep and defect: bit-identical to the oracle in FP64 and FP32 (same
operation order, no contraction), on blocks of edge 16, 32 and 64 and on the
full 512^3 grid of 64^3 blocks (tanh-stretched z, Dirichlet 0), and with
periodic, zero-gradient, mirrored and Dirichlet-constant faces on blocks
whose spacings differ per direction.  Refined solve (FP32 sweeps, FP64 defect
and correction) to ||r||/||f|| <= 1e-12 against the direct solution of the
independently assembled sparse matrix and against an all-FP64 Jacobi solve."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import cubismnova_amd as cnv
import oracle
from cubismnova_amd import stretch
from cubismnova_amd.layout import blocks_to_global, global_to_blocks
from tests.helpers import assert_same, face_bc_lists
from tests.test_oracle_jacobi import D0, matrix

pytestmark = pytest.mark.gpu


def grid(nb, n, dtype, zf, ncomp=1):
    g = cnv.Cartesian(nb, n, dtype=dtype, ncomp=ncomp)
    for d in range(3):
        for s in (0, 1):
            g.addFaceBC(cnv.Dirichlet(d, s, 0.0))
    cnv.set_zfaces(g, zf)
    return g


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("edge", [16, 32, 64])
def test_sweep_and_defect_bitexact(gpu, dtype, edge):
    nb, n = (2, 2, 2), (edge,) * 3
    N = 2 * edge
    zf = stretch.tanh_faces(N, 2.0)
    czm, czp, _, _ = stretch.metric(zf)
    h = [1.0 / N] * 3
    rng = np.random.Generator(np.random.PCG64(edge))
    u0 = rng.standard_normal((N, N, N)).astype(dtype)
    f0 = rng.standard_normal((N, N, N)).astype(dtype)
    u, f, out = grid(nb, n, dtype, zf), grid(nb, n, dtype, zf), grid(nb, n, dtype, zf)
    u.upload_global(0, u0)
    f.upload_global(0, f0)
    fo = oracle.Field(global_to_blocks(u0, nb, n), nb, n, bcs=face_bc_lists(nb, D0))
    fs = global_to_blocks(f0, nb, n)
    cnv.jacobi_sweep(u, f, out, 2.0 / 3.0)
    assert_same(out.download(0), oracle.jacobi(fo, fs, czm, czp, h, 2.0 / 3.0, 0), "sweep")
    s = cnv.DeviceBuffer(8)
    cnv.jacobi_defect(u, f, out, sumsq=s)
    r, sref = oracle.jacobi(fo, fs, czm, czp, h, 2.0 / 3.0, 1)
    assert_same(out.download(0), r, "defect")
    assert abs(s.download(np.float64, 1)[0] - sref) <= 1e-12 * sref
    if dtype == np.float64:
        r32 = grid(nb, n, np.float32, zf)
        cnv.jacobi_defect(u, f, r32, sumsq=s)
        assert_same(r32.download(0), r.astype(np.float32), "defect rounded to FP32")


def test_refined_solve_matches_direct_and_fp64_jacobi(gpu):
    nb, n, N = (2, 2, 2), (16, 16, 16), 32
    zf = stretch.tanh_faces(N, 2.0)
    czm, czp, dz, zc = stretch.metric(zf)
    h = [1.0 / N] * 3
    x = (np.arange(N) + 0.5) / N
    src = np.zeros((N, N, N))
    rng = np.random.Generator(np.random.PCG64(77))
    for _ in range(3):
        c = rng.uniform(0.2, 0.8, 3)
        amp = rng.uniform(-1, 1)
        src += amp * np.exp(-((x[None, None, :] - c[0]) ** 2 + (x[None, :, None] - c[1]) ** 2
                              + (zc[:, None, None] - c[2]) ** 2) / (2 * 0.05 ** 2))
    src -= src.mean()
    exact = spla.spsolve(matrix(N, czm, czp, h).tocsc(), src.ravel()).reshape(N, N, N)

    u = grid(nb, n, np.float64, zf)
    f = grid(nb, n, np.float64, zf)
    u.upload_global(0, np.zeros((N, N, N)))
    f.upload_global(0, src)
    work = grid(nb, n, np.float32, zf, ncomp=3)
    cyc, sw, rel, hist = cnv.jacobi_solve(u, f, work, refine=True, sweeps_per_cycle=50, max_cycles=4000, tol=1e-12)
    assert rel <= 1e-12, (cyc, rel)
    assert np.all(np.diff(hist[: max(2, len(hist) // 2)]) < 0)
    ur = u.download_global(0)

    u2 = grid(nb, n, np.float64, zf)
    u2.upload_global(0, np.zeros((N, N, N)))
    work2 = grid(nb, n, np.float64, zf, ncomp=2)
    cyc2, sw2, rel2, _ = cnv.jacobi_solve(u2, f, work2, refine=False, sweeps_per_cycle=50, max_cycles=4000,
                                          tol=1e-12)
    assert rel2 <= 1e-12, (cyc2, rel2)
    u64 = u2.download_global(0)
    scale = np.max(np.abs(exact))
    assert np.max(np.abs(ur - exact)) <= 1e-9 * scale
    assert np.max(np.abs(u64 - exact)) <= 1e-9 * scale
    assert np.max(np.abs(ur - u64)) <= 1e-9 * scale
    # the refined solve needs about as many sweeps as FP64 Jacobi
    assert sw <= 1.2 * sw2 + 100


def test_solver_rejects_bad_options(gpu):
    nb, n = (1, 1, 1), (16, 16, 16)
    zf = stretch.uniform_faces(16)
    u, f = grid(nb, n, np.float64, zf), grid(nb, n, np.float64, zf)
    with pytest.raises(cnv.CnvError):
        cnv.jacobi_solve(u, f, grid(nb, n, np.float64, zf, ncomp=3), refine=True)  # work must be FP32
    with pytest.raises(cnv.CnvError):
        cnv.jacobi_solve(u, f, grid(nb, n, np.float32, zf, ncomp=3), sweeps_per_cycle=3)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_full_size_sweep_512_bitexact(gpu, dtype):
    """The 512^3 stretched-z Poisson grid of 8x8x8 blocks of 64^3 (Dirichlet 0
    on every face): one sweep and one defect, every cell bit-exact against
    the oracle run on the same bytes."""
    from cubismnova_amd import workloads
    N, nb, n = 512, (8, 8, 8), (64, 64, 64)
    zf = stretch.tanh_faces(N, 2.0)
    czm, czp, _, zc = stretch.metric(zf)
    h = [1.0 / N] * 3
    rng = np.random.Generator(np.random.PCG64(512))
    u0 = rng.random((N, N, N), dtype=np.float32 if dtype == np.float32 else np.float64).astype(dtype)
    f0 = workloads.poisson_source(N, zc).astype(dtype)
    u, f, out = grid(nb, n, dtype, zf), grid(nb, n, dtype, zf), grid(nb, n, dtype, zf)
    u.upload_global(0, u0)
    f.upload_global(0, f0)
    fo = oracle.Field(global_to_blocks(u0, nb, n), nb, n, bcs=face_bc_lists(nb, D0))
    fs = global_to_blocks(f0, nb, n)
    del u0, f0
    nt = oracle.max_threads()
    cnv.jacobi_sweep(u, f, out, 2.0 / 3.0)
    ref = out.download(0)
    assert_same(ref, oracle.jacobi(fo, fs, czm, czp, h, 2.0 / 3.0, 0, nthreads=nt), "sweep 512^3")
    s = cnv.DeviceBuffer(8)
    cnv.jacobi_defect(u, f, out, sumsq=s)
    r, sref = oracle.jacobi(fo, fs, czm, czp, h, 2.0 / 3.0, 1, nthreads=nt)
    assert_same(out.download(0), r, "defect 512^3")
    assert abs(s.download(np.float64, 1)[0] - sref) <= 1e-12 * sref
    # every sweep tile the autotuner can pick, at full size (several rounds
    # of workgroups): bit-identical to the sweep checked above
    saved = cnv.tune_jacobi(dtype, 64)
    try:
        ran = 0
        for wg, zc, pf in sweep_tiles(64):
            try:
                cnv.set_tune_jacobi(dtype, 64, wg, zc, pf)
            except cnv.CnvError:
                continue
            clear(out)
            cnv.jacobi_sweep(u, f, out, 2.0 / 3.0)
            assert_same(out.download(0), ref, "sweep 512^3, tile %d x %d, prefetch %d" % (wg, zc, pf))
            ran += 1
        assert ran >= 12
    finally:
        cnv.set_tune_jacobi(dtype, 64, *saved)


def clear(g):
    """component 0 of g set to zero (a sweep that wrote nothing would fail)"""
    code = 1 if g.dtype == np.float64 else 0
    cnv.call("cnv_fill", g.data_ptr(0), code, 0.0, g.component_bytes // g.dtype.itemsize, None)


def sweep_tiles(edge):
    """the autotuner's candidates for blocks of this edge (z chunks up to the edge)"""
    return [(wg, zc, pf) for wg in (64, 128, 256, 512) for zc in (8, 16, 32, 64) if zc <= edge for pf in (0, 1)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("edge", [16, 32, 64])
def test_every_sweep_tile_bitexact(gpu, dtype, edge):
    """Every (workgroup size, z-planes per tile, f prefetch) sweep tile of
    the autotuner bit-exact against the oracle on 2x2x2 blocks, then the
    autotuner itself: it picks one of them and leaves the sweep exact."""
    nb, n = (2, 2, 2), (edge,) * 3
    N = 2 * edge
    zf = stretch.tanh_faces(N, 2.0)
    czm, czp, _, _ = stretch.metric(zf)
    h = [1.0 / N] * 3
    rng = np.random.Generator(np.random.PCG64(7 * edge))
    u0 = rng.standard_normal((N, N, N)).astype(dtype)
    f0 = rng.standard_normal((N, N, N)).astype(dtype)
    u, f, out = grid(nb, n, dtype, zf), grid(nb, n, dtype, zf), grid(nb, n, dtype, zf)
    u.upload_global(0, u0)
    f.upload_global(0, f0)
    fo = oracle.Field(global_to_blocks(u0, nb, n), nb, n, bcs=face_bc_lists(nb, D0))
    ref = oracle.jacobi(fo, global_to_blocks(f0, nb, n), czm, czp, h, 2.0 / 3.0, 0)
    saved = cnv.tune_jacobi(dtype, edge)
    try:
        ran = 0
        for wg, zc, pf in sweep_tiles(edge):
            try:
                cnv.set_tune_jacobi(dtype, edge, wg, zc, pf)
            except cnv.CnvError:
                continue
            clear(out)
            cnv.jacobi_sweep(u, f, out, 2.0 / 3.0)
            assert_same(out.download(0), ref, "tile %d x %d, prefetch %d" % (wg, zc, pf))
            ran += 1
        assert ran >= 4
        wg, zc, pf, table = cnv.autotune_jacobi(u, f, out, iters=2)
        assert (wg, zc, pf) in [t[:3] for t in table] and cnv.tune_jacobi(dtype, edge) == (wg, zc, pf)
        assert len(table) == ran and all(t[3] > 0 for t in table)
        clear(out)
        cnv.jacobi_sweep(u, f, out, 2.0 / 3.0)
        assert_same(out.download(0), ref, "autotuned tile")
    finally:
        cnv.set_tune_jacobi(dtype, edge, *saved)
    with pytest.raises(cnv.CnvError):
        cnv.set_tune_jacobi(dtype, edge, 96, 16, 0)


def test_refined_solve_on_64_blocks(gpu):
    """The refined solver on blocks of edge 64 (the edge of the 512^3
    workload), one block of 64^3: FP32 sweeps with FP64 defect and
    correction reach ||r||/||f|| <= 1e-12 and agree with all-FP64 Jacobi."""
    from cubismnova_amd import workloads
    nb, n, N = (1, 1, 1), (64, 64, 64), 64
    zf = stretch.tanh_faces(N, 2.0)
    _, _, _, zc = stretch.metric(zf)
    src = workloads.poisson_source(N, zc)
    u = grid(nb, n, np.float64, zf)
    f = grid(nb, n, np.float64, zf)
    u.upload_global(0, np.zeros((N, N, N)))
    f.upload_global(0, src)
    cyc, sw, rel, hist = cnv.jacobi_solve(u, f, grid(nb, n, np.float32, zf, ncomp=3), refine=True,
                                          sweeps_per_cycle=50, max_cycles=4000, tol=1e-12)
    assert rel <= 1e-12, (cyc, rel)
    u2 = grid(nb, n, np.float64, zf)
    u2.upload_global(0, np.zeros((N, N, N)))
    cyc2, sw2, rel2, _ = cnv.jacobi_solve(u2, f, grid(nb, n, np.float64, zf, ncomp=2), refine=False,
                                          sweeps_per_cycle=50, max_cycles=4000, tol=1e-12)
    assert rel2 <= 1e-12, (cyc2, rel2)
    a, b = u.download_global(0), u2.download_global(0)
    assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b))
    assert sw <= 1.2 * sw2 + 100


def test_capture_after_grids_were_freed(gpu):
    """Sweeps recorded into a HIP graph in turn, each time on grids made after
    the previous grids, graph and stream were destroyed in the same process:
    every capture succeeds and every replay equals the same sweeps run
    directly."""
    nb, n, N = (2, 2, 2), (64, 64, 64), 128
    zf = stretch.tanh_faces(N, 2.0)
    rng = np.random.Generator(np.random.PCG64(11))
    u0 = rng.standard_normal((N, N, N)).astype(np.float32)
    f0 = rng.standard_normal((N, N, N)).astype(np.float32)
    saved = cnv.tune_jacobi(np.float32, 64)
    try:
        for rnd, tile in enumerate([(256, 64, 0), (512, 64, 0), (256, 64, 0)]):
            cnv.set_tune_jacobi(np.float32, 64, *tile)
            a, b, f = (grid(nb, n, np.float32, zf) for _ in range(3))
            f.upload_global(0, f0)
            for g in (a, b, f):
                g.prepare()
            a.upload_global(0, u0)
            for k in range(4):
                x, y = (a, b) if k % 2 == 0 else (b, a)
                cnv.jacobi_sweep(x, f, y, 2.0 / 3.0)
            want = a.download(0)
            st = cnv.Stream()
            gr = cnv.Graph(st).capture()
            for k in range(4):
                x, y = (a, b) if k % 2 == 0 else (b, a)
                cnv.jacobi_sweep(x, f, y, 2.0 / 3.0, stream=st)
            gr.end()
            for rep in range(2):
                a.upload_global(0, u0)
                gr.launch()
                st.synchronize()
                assert_same(a.download(0), want, "round %d tile %s replay %d" % (rnd, tile, rep))
            del gr, a, b, f, st
    finally:
        cnv.set_tune_jacobi(np.float32, 64, *saved)


# face BC sets as (kind, dir, side, value): 1 absorbing, 2 Dirichlet, 3 symmetry (value = sign)
BC_SETS = {
    "periodic": [],
    "mixed": [(1, 1, 0, 0.0), (1, 1, 1, 0.0), (2, 2, 0, 0.5), (3, 2, 1, -1.0)],
    "symmetric": [(3, d, s, 1.0) for d in range(3) for s in (0, 1)],
    "dirichlet": [(2, d, s, 0.25 * (1 + d) * (1 if s else -1)) for d in range(3) for s in (0, 1)],
}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("bcs", list(BC_SETS), ids=list(BC_SETS))
def test_sweep_and_defect_other_bcs(gpu, dtype, bcs):
    """Sweep and defect bit-exact against the oracle with face BCs other than
    Dirichlet 0 (periodic, zero-gradient, symmetric and antisymmetric
    mirrors, Dirichlet constants) on 3 x 2 x 2 blocks of 16^3, whose x, y
    and z spacings differ."""
    nb, n = (3, 2, 2), (16, 16, 16)
    NX, NY, NZ = 48, 32, 32
    zf = stretch.tanh_faces(NZ, 2.0)
    czm, czp, _, _ = stretch.metric(zf)
    h = [1.0 / NX, 1.0 / NY, 1.0 / NZ]
    make = {1: lambda d, s, v: cnv.Absorbing(d, s), 2: lambda d, s, v: cnv.Dirichlet(d, s, v),
            3: lambda d, s, v: cnv.Symmetry(d, s, v)}

    def bgrid():
        g = cnv.Cartesian(nb, n, dtype=dtype)
        for (k, d, s, v) in BC_SETS[bcs]:
            g.addFaceBC(make[k](d, s, v))
        cnv.set_zfaces(g, zf)
        return g

    rng = np.random.Generator(np.random.PCG64(len(bcs) + 31))
    u0 = rng.standard_normal((NZ, NY, NX)).astype(dtype)
    f0 = rng.standard_normal((NZ, NY, NX)).astype(dtype)
    u, f, out = bgrid(), bgrid(), bgrid()
    u.upload_global(0, u0)
    f.upload_global(0, f0)
    fo = oracle.Field(global_to_blocks(u0, nb, n), nb, n, bcs=face_bc_lists(nb, BC_SETS[bcs]))
    fs = global_to_blocks(f0, nb, n)
    cnv.jacobi_sweep(u, f, out, 2.0 / 3.0)
    assert_same(out.download(0), oracle.jacobi(fo, fs, czm, czp, h, 2.0 / 3.0, 0), "sweep")
    s = cnv.DeviceBuffer(8)
    cnv.jacobi_defect(u, f, out, sumsq=s)
    r, sref = oracle.jacobi(fo, fs, czm, czp, h, 2.0 / 3.0, 1)
    assert_same(out.download(0), r, "defect")
    assert abs(s.download(np.float64, 1)[0] - sref) <= 1e-12 * sref
