"""GPU parity of the 7-point Laplacian and the fused diffusion step against
the oracle.  Every comparison is bit-exact (same operation order, no FMA
contraction) except FP64 reductions, which are order-dependent and checked to
1e-13 relative against the oracle's long-double sum."""
import numpy as np
import pytest

import cubismnova_amd as cnv
import oracle
from cubismnova_amd import workloads
from tests.helpers import assert_same, cells_of, face_bc_lists, random_slab

pytestmark = pytest.mark.gpu

KIND = {0: cnv.Base, 1: cnv.Absorbing, 2: cnv.Dirichlet, 3: cnv.Symmetry}


def mk_bc(t):
    k, d, s, v = t
    return KIND[k](d, s, v) if k in (2, 3) else KIND[k](d, s)


def setup(nb, n, dtype, slab, bcl=None, ncomp=1):
    g = cnv.Cartesian(nb, n, dtype=dtype, ncomp=ncomp)
    g.upload(0, slab)
    for b, lst in (bcl or {}).items():
        bi = (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
        g.setBC(bi, [mk_bc(t) for t in lst])
    return g


def test_small_config_laplacian_and_euler(gpu):
    """4^3 blocks of 16^3, FP64, periodic: ghosts, Laplacian and one Euler step."""
    nb, n, N = (4, 4, 4), (16, 16, 16), 64
    u0 = workloads.diffusion_small(N)
    g = cnv.Cartesian(nb, n)
    g.upload_global(0, u0)
    slab = g.download(0)
    f = oracle.Field(slab, nb, n)
    # ghosts of every block
    buf, ext = g.lab_all(cnv.Stencil(-1, 2))
    buf.fill(np.float64, np.nan)
    g.lab_all(cnv.Stencil(-1, 2), out=buf)
    labs = buf.download(np.float64).reshape(64, ext[2], ext[1], ext[0])
    for b in range(64):
        bi = (b % 4, (b // 4) % 4, b // 16)
        want = oracle.lab_load(f, bi, -1, 2)
        m = ~np.isnan(want)
        assert np.array_equal(labs[b][m], want[m])
    h = [1.0 / 64] * 3
    lap = cnv.Cartesian(nb, n)
    cnv.laplacian(g, lap)
    assert np.array_equal(lap.download(0), oracle.laplacian(f, h))
    dt = 0.1 * h[0] ** 2
    u1 = cnv.Cartesian(nb, n)
    lap2 = cnv.Cartesian(nb, n)
    s = cnv.DeviceBuffer(8)
    cnv.diffusion_step(g, u1, dt, lap=lap2, sumsq=s)
    un, sref, lref = oracle.diffusion_step(f, h, dt, want_lap=True)
    assert np.array_equal(lap2.download(0), lref)
    got = u1.download(0)
    assert np.array_equal(got, un)
    rel = np.max(np.abs(got - un)) / np.max(np.abs(un))
    assert rel <= 1e-14
    ssum = s.download(np.float64, 1)[0]
    assert abs(ssum - sref) <= 1e-13 * sref


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("edge", [16, 32, 64])
@pytest.mark.parametrize("bcs", ["periodic", "mixed"])
def test_diffusion_specialised_kernels(gpu, dtype, edge, bcs):
    nb = (2, 2, 2) if edge < 64 else (2, 1, 1)
    n = (edge,) * 3
    slab = random_slab(31 + edge, nb, n, dtype)
    bcl = None
    if bcs == "mixed":
        bcl = face_bc_lists(nb, [(2, 0, 0, 0.5), (1, 0, 1, 0.0), (3, 1, 0, -1.0), (1, 1, 1, 0.0),
                                 (2, 2, 0, -0.25), (3, 2, 1, 1.0)])
    g = setup(nb, n, dtype, slab, bcl)
    f = oracle.Field(slab, nb, n, bcs=bcl)
    h = [1.0 / (nb[d] * edge) for d in range(3)]
    out = cnv.Cartesian(nb, n, dtype=dtype)
    s = cnv.DeviceBuffer(8)
    cnv.diffusion_step(g, out, 0.1 * h[0] ** 2, sumsq=s)
    un, sref = oracle.diffusion_step(f, h, 0.1 * h[0] ** 2)
    assert_same(out.download(0), un, "euler")
    assert abs(s.download(np.float64, 1)[0] - sref) <= 1e-13 * sref
    lap = cnv.Cartesian(nb, n, dtype=dtype)
    cnv.laplacian(g, lap)
    assert_same(lap.download(0), oracle.laplacian(f, h), "laplacian")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_diffusion_generic_shapes(gpu, dtype):
    for nb, n in [((3, 2, 2), (12, 8, 10)), ((1, 1, 1), (5, 7, 3)), ((2, 2, 2), (8, 8, 8))]:
        bs = cnv.Cartesian(nb, n, dtype=dtype).block_stride   # padded to 32 bytes per block
        slab = random_slab(5, nb, n, dtype, block_stride=bs)
        bcl = face_bc_lists(nb, [(2, 1, 0, 1.0), (1, 2, 1, 0.0)])
        g = setup(nb, n, dtype, slab, bcl)
        f = oracle.Field(slab, nb, n, bs, bcs=bcl)
        h = [1.0 / (nb[d] * n[d]) for d in range(3)]
        out = cnv.Cartesian(nb, n, dtype=dtype)
        s = cnv.DeviceBuffer(8)
        cnv.diffusion_step(g, out, 1e-6, sumsq=s)
        un, sref = oracle.diffusion_step(f, h, 1e-6)
        assert_same(cells_of(out.download(0), nb, n, bs), cells_of(un, nb, n, bs), str(n))
        assert abs(s.download(np.float64, 1)[0] - sref) <= 1e-13 * sref


def test_diffusion_multicomponent(gpu):
    nb, n = (2, 2, 2), (16, 16, 16)
    g = cnv.Cartesian(nb, n, ncomp=3)
    out = cnv.Cartesian(nb, n, ncomp=3)
    h = [1.0 / 32] * 3
    refs, tot = [], 0.0
    for c in range(3):
        slab = random_slab(40 + c, nb, n)
        g.upload(c, slab)
        un, s = oracle.diffusion_step(oracle.Field(slab, nb, n), h, 1e-5)
        refs.append(un)
        tot += s
    sbuf = cnv.DeviceBuffer(8)
    cnv.diffusion_step(g, out, 1e-5, sumsq=sbuf)
    for c in range(3):
        assert np.array_equal(out.download(c), refs[c])
    assert abs(sbuf.download(np.float64, 1)[0] - tot) <= 1e-13 * tot


def test_argument_errors(gpu):
    g = cnv.Cartesian(1, 8)
    out = cnv.Cartesian(1, 8)
    g.setBC((0, 0, 0), [cnv.Absorbing(0, 0), cnv.Base(0, 1)])  # last of dir 0 periodic: x- absorbing
    cnv.diffusion_step(g, out, 1e-3)
    with pytest.raises(cnv.CnvError):
        cnv.diffusion_step(g, g, 1e-3)             # in-place is rejected
    with pytest.raises(cnv.CnvError):
        cnv.diffusion_step(g, cnv.Cartesian(1, 16), 1e-3)   # shape mismatch
    with pytest.raises(cnv.CnvError):
        cnv.diffusion_step(g, cnv.Cartesian(1, 8, dtype=np.float32), 1e-3)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_full_size_100_steps(gpu, dtype):
    """8^3 blocks of 32^3 (256^3), 100 Euler steps with an FP64 norm every step,
    against the oracle run on the same bytes: final field bit-exact, norms 1e-13."""
    nb, n, N = (8, 8, 8), (32, 32, 32), 256
    u0 = workloads.diffusion_large(N).astype(dtype)
    a = cnv.Cartesian(nb, n, dtype=dtype)
    b = cnv.Cartesian(nb, n, dtype=dtype)
    a.upload_global(0, u0)
    slab = a.download(0)
    h = [1.0 / N] * 3
    dt = 0.1 * h[0] ** 2
    norms = cnv.DeviceBuffer(8 * 100)
    st = cnv.Stream()
    src, dst = a, b
    for k in range(100):
        cnv.diffusion_step(src, dst, dt, sumsq=norms.at(8 * k), stream=st)
        src, dst = dst, src
    st.synchronize()
    got = src.download(0)
    gn = norms.download(np.float64, 100)
    nt = max(1, min(16, oracle.max_threads()))
    cur = slab
    for k in range(100):
        cur, s = oracle.diffusion_step(oracle.Field(cur, nb, n), h, dt, nthreads=nt)
        assert abs(gn[k] - s) <= 1e-13 * s, k
    assert_same(got, cur, "256^3 after 100 steps")
    # conservation of the periodic mean
    assert abs(got.astype(np.float64).sum() - slab.astype(np.float64).sum()) <= 1e-9 * np.abs(slab).sum()


def test_fp32_vs_fp64_error(gpu):
    nb, n, N = (8, 8, 8), (32, 32, 32), 256
    u0 = workloads.diffusion_large(N)
    h = 1.0 / N
    dt = 0.1 * h * h
    res = {}
    for dtype in (np.float64, np.float32):
        a = cnv.Cartesian(nb, n, dtype=dtype)
        b = cnv.Cartesian(nb, n, dtype=dtype)
        a.upload_global(0, u0.astype(dtype))
        src, dst = a, b
        for k in range(100):
            cnv.diffusion_step(src, dst, dt)
            src, dst = dst, src
        res[dtype] = src.download(0).astype(np.float64)
    rel = np.linalg.norm(res[np.float32] - res[np.float64]) / np.linalg.norm(res[np.float64])
    assert rel <= 1e-6, rel


def test_sums_deterministic(gpu):
    nb, n = (4, 4, 4), (16, 16, 16)
    slab = random_slab(77, nb, n)
    g = setup(nb, n, np.float64, slab)
    a = cnv.grid_sums(g)
    b = cnv.grid_sums(g)
    assert a == b
    assert abs(a[0] - slab.sum()) <= 1e-12 * np.abs(slab).sum()
    assert abs(a[1] - (slab * slab).sum()) <= 1e-13 * (slab * slab).sum()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nsteps", [1, 2, 7])
def test_run_matches_steps(gpu, dtype, nsteps):
    """cnv_diffusion_run: the same fields as nsteps single steps (bit-exact) and
    norms (reduced inside the next step's launch) against the oracle to 1e-13,
    on a specialised shape with mixed BCs and on a generic shape."""
    for nb, n in [((2, 2, 2), (16, 16, 16)), ((3, 2, 1), (12, 8, 10))]:
        bs = cnv.Cartesian(nb, n, dtype=dtype).block_stride
        slab = random_slab(90 + nsteps, nb, n, dtype, block_stride=bs)
        bcl = face_bc_lists(nb, [(2, 0, 0, 0.5), (1, 0, 1, 0.0), (3, 2, 1, -1.0)])
        u = setup(nb, n, dtype, slab, bcl)
        tmp = setup(nb, n, dtype, slab, bcl)
        h = [1.0 / (nb[d] * n[d]) for d in range(3)]
        dt = 0.1 * min(h) ** 2
        norms = cnv.DeviceBuffer(8 * nsteps)
        norms.fill(np.float64, np.nan)
        res = cnv.diffusion_run(u, tmp, nsteps, dt, norms=norms)
        assert res is (u if nsteps % 2 == 0 else tmp)
        got = res.download(0)
        gn = norms.download(np.float64, nsteps)
        cur = slab
        for k in range(nsteps):
            cur, s = oracle.diffusion_step(oracle.Field(cur, nb, n, bs, bcs=bcl), h, dt)
            assert abs(gn[k] - s) <= 1e-13 * s, (n, k, gn[k], s)
        assert_same(cells_of(got, nb, n, bs), cells_of(cur, nb, n, bs), "run %s" % (n,))
        # without norms
        u.upload(0, slab)
        res = cnv.diffusion_run(u, tmp, nsteps, dt)
        assert np.array_equal(cells_of(res.download(0), nb, n, bs), cells_of(got, nb, n, bs))


def test_run_full_size_norm_order(gpu):
    """256^3 FP64 run of 100 steps against 100 single steps: identical fields,
    norms equal to 1e-14 (the two paths sum the same partials in different
    fixed trees) and the same on a repeat (deterministic)."""
    nb, n, N = (8, 8, 8), (32, 32, 32), 256
    u0 = workloads.diffusion_large(N)
    h = 1.0 / N
    dt = 0.1 * h * h
    a, b = cnv.Cartesian(nb, n), cnv.Cartesian(nb, n)
    a.upload_global(0, u0)
    ns = cnv.DeviceBuffer(8 * 100)
    src, dst = a, b
    for k in range(100):
        cnv.diffusion_step(src, dst, dt, sumsq=ns.at(8 * k))
        src, dst = dst, src
    want = src.download(0)
    wn = ns.download(np.float64, 100)
    runs = []
    for rep in range(2):
        a.upload_global(0, u0)
        nr = cnv.DeviceBuffer(8 * 100)
        res = cnv.diffusion_run(a, b, 100, dt, norms=nr)
        assert np.array_equal(res.download(0), want)
        runs.append(nr.download(np.float64, 100))
    assert np.array_equal(runs[0], runs[1])
    assert np.max(np.abs(runs[0] - wn) / wn) <= 1e-14


def test_run_argument_errors(gpu):
    u, t = cnv.Cartesian(2, 8), cnv.Cartesian(2, 8)
    u.setBC((0, 0, 0), [cnv.Absorbing(0, 0)])
    with pytest.raises(cnv.CnvError):
        cnv.diffusion_run(u, t, 2, 1e-3)            # BC lists differ
    cnv.diffusion_run(u, t, 1, 1e-3)                # one step reads only u
    with pytest.raises(cnv.CnvError):
        cnv.diffusion_run(u, u, 2, 1e-3)
    with pytest.raises(cnv.CnvError):
        cnv.diffusion_run(u, t, -1, 1e-3)
    cnv.diffusion_run(t, cnv.Cartesian(2, 8), 0, 1e-3)


TWO_STEP_SHAPES = {np.float32: [(16, 64), (16, 128), (32, 128), (32, 256), (32, 512)],
                   np.float64: [(16, 64), (16, 128), (32, 256), (32, 512)]}
# the library's built-in (workgroup size, z depth) per block edge
BUILTIN_TWO_STEP = {np.float32: {16: (64, 16), 32: (256, 16)}, np.float64: {16: (128, 16), 32: (256, 32)}}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_two_step_kernel_matches_single_steps(gpu, dtype):
    """Periodic grids take the two-steps-per-launch kernel inside diffusion_run:
    fields bit-identical to single steps and norms within 1e-13 of the oracle,
    for every tile shape and z depth (8, 16 and 32 planes), odd and even step counts, one-block
    directions (self-periodic) and several blocks."""
    for nb, e in [((2, 2, 2), 16), ((3, 2, 1), 32), ((1, 1, 1), 16), ((2, 1, 2), 32)]:
        n = (e, e, e)
        slab = random_slab(300 + e + nb[0], nb, n, dtype)
        h = [1.0 / (nb[d] * e) for d in range(3)]
        dt = 0.1 * min(h) ** 2
        nsteps = 7
        ref = [slab]
        refn = []
        for k in range(nsteps):
            un, s = oracle.diffusion_step(oracle.Field(ref[-1], nb, n), h, dt)
            ref.append(un)
            refn.append(s)
        for e2, wg in TWO_STEP_SHAPES[dtype]:
            if e2 != e:
                continue
            for z in sorted({8, 16, e}):
                cnv.set_tune_diffusion_run(dtype, e, wg, z)
                for ns in (4, 5, nsteps):
                    u = cnv.Cartesian(nb, n, dtype=dtype)
                    t = cnv.Cartesian(nb, n, dtype=dtype)
                    u.upload(0, slab)
                    norms = cnv.DeviceBuffer(8 * ns)
                    norms.fill(np.float64, np.nan)
                    res = cnv.diffusion_run(u, t, ns, dt, norms=norms)
                    assert_same(res.download(0), ref[ns], "%s wg %d z %d nsteps %d" % (nb, wg, z, ns))
                    gn = norms.download(np.float64, ns)
                    for k in range(ns):
                        assert abs(gn[k] - refn[k]) <= 1e-13 * refn[k], (nb, wg, z, ns, k)
    # back to the built-in choices for the tests that follow
    cnv.set_tune_diffusion_run(dtype, 16, *BUILTIN_TWO_STEP[dtype][16])
    cnv.set_tune_diffusion_run(dtype, 32, *BUILTIN_TWO_STEP[dtype][32])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("edge,N,steps", [(32, 256, 100), (16, 256, 100), (32, 512, 20)],
                         ids=["32^3-blocks-256", "16^3-blocks-256", "32^3-blocks-512"])
def test_two_step_full_size_graph_every_tile(gpu, dtype, edge, N, steps):
    """Full-size grids (256^3 of 32^3 or 16^3 blocks, and 512^3 of 32^3
    blocks: several rounds of workgroups per launch for every tile), `steps`
    steps through diffusion_run for every tile shape of the block edge, run
    directly and as a captured HIP graph replayed twice (the benchmark's use):
    fields bit-identical to single steps each time."""
    n = (edge,) * 3
    nb = (N // edge,) * 3
    u0 = workloads.diffusion_large(N).astype(dtype)
    dt = 0.1 / N / N
    a, b = cnv.Cartesian(nb, n, dtype=dtype), cnv.Cartesian(nb, n, dtype=dtype)
    a.upload_global(0, u0)
    src, dst = a, b
    for k in range(steps):
        cnv.diffusion_step(src, dst, dt)
        src, dst = dst, src
    want = src.download(0)
    st = cnv.Stream()
    a.prepare()
    b.prepare()
    try:
        for e2, wg in TWO_STEP_SHAPES[dtype]:
            if e2 != edge:
                continue
            for z in sorted({8, 16, edge}):
                cnv.set_tune_diffusion_run(dtype, edge, wg, z)
                a.upload_global(0, u0)
                res = cnv.diffusion_run(a, b, steps, dt)
                assert_same(res.download(0), want, "direct wg %d z %d" % (wg, z))
                g = cnv.Graph(st).capture()
                res = cnv.diffusion_run(a, b, steps, dt, stream=st)
                g.end()
                for rep in range(2):
                    a.upload_global(0, u0)
                    g.launch()
                    st.synchronize()
                    assert_same(res.download(0), want, "graph wg %d z %d replay %d" % (wg, z, rep))
    finally:
        cnv.set_tune_diffusion_run(dtype, edge, *BUILTIN_TWO_STEP[dtype][edge])


def test_two_step_autotune_and_fallback(gpu):
    nb, n = (2, 2, 2), (32, 32, 32)
    u = cnv.Cartesian(nb, n, dtype=np.float32)
    t = cnv.Cartesian(nb, n, dtype=np.float32)
    u.upload(0, random_slab(5, nb, n, np.float32))
    wg, z, table = cnv.autotune_diffusion_run(u, t, iters=2)
    assert (wg, z) == cnv.tune_diffusion_run(np.float32, 32)
    assert len(table) >= 3 and all(r[2] > 0 for r in table)
    # a grid with a BC list cannot use the two-step kernel
    u.addFaceBC(cnv.Absorbing(0, 0))
    with pytest.raises(cnv.CnvError):
        cnv.autotune_diffusion_run(u, t, iters=1)
    with pytest.raises(cnv.CnvError):
        cnv.set_tune_diffusion_run(np.float32, 32, 64, 8)     # shape that runs out of registers


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_two_step_kernel_multicomponent(gpu, dtype):
    """Three components in one periodic grid through diffusion_run (two steps
    per launch): each component bit-identical to its own single steps, the
    norms summed over components."""
    nb, n = (2, 3, 2), (16, 16, 16)
    u = cnv.Cartesian(nb, n, dtype=dtype, ncomp=3)
    t = cnv.Cartesian(nb, n, dtype=dtype, ncomp=3)
    h = [1.0 / (nb[d] * 16) for d in range(3)]
    dt = 0.1 * min(h) ** 2
    nsteps = 6
    refs, tot = [], np.zeros(nsteps)
    for c in range(3):
        slab = random_slab(500 + c, nb, n, dtype)
        u.upload(c, slab)
        cur = slab
        for k in range(nsteps):
            cur, s = oracle.diffusion_step(oracle.Field(cur, nb, n), h, dt)
            tot[k] += s
        refs.append(cur)
    norms = cnv.DeviceBuffer(8 * nsteps)
    res = cnv.diffusion_run(u, t, nsteps, dt, norms=norms)
    assert res is u
    for c in range(3):
        assert_same(res.download(c), refs[c], "component %d" % c)
    gn = norms.download(np.float64, nsteps)
    assert np.all(np.abs(gn - tot) <= 1e-13 * tot), (gn, tot)


def test_two_gib_component_slab(gpu):
    """A grid whose component slab is 2^31 bytes (1024 x 1024 x 512 FP32
    cells in 32^3 blocks): the two-step kernel, which addresses a slab with
    32-bit byte offsets, is not used and diffusion_run takes single steps.
    The field is a separable Fourier mode, an eigenfunction of the discrete
    7-point Laplacian on the periodic grid, so one step multiplies it by
    1 + dt * lambda; every cell, the highest addresses included, is checked
    against that after one step and after a two-step run."""
    nb, n = (32, 32, 16), (32, 32, 32)
    NX, NY, NZ = nb[0] * n[0], nb[1] * n[1], nb[2] * n[2]
    k = (4, 4, 2)
    h = (1.0 / NX, 1.0 / NY, 1.0 / NZ)
    lam = sum((2 * np.cos(2 * np.pi * k[d] * h[d]) - 2) / h[d] ** 2 for d in range(3))
    dt = 0.1 * h[0] ** 2
    rows = [np.cos(2 * np.pi * k[d] * (np.arange(N) + 0.5) * h[d]) for d, N in enumerate((NX, NY, NZ))]
    u0 = (rows[2][:, None, None].astype(np.float32) * rows[1][None, :, None].astype(np.float32)
          * rows[0][None, None, :].astype(np.float32))
    a = cnv.Cartesian(nb, n, dtype=np.float32)
    b = cnv.Cartesian(nb, n, dtype=np.float32)
    assert a.component_bytes == 1 << 31
    a.upload_global(0, u0)

    def check(res, steps):
        got = res.download_global(0)
        f = (1.0 + dt * lam) ** steps
        for z0 in range(0, NZ, 64):
            d = np.abs(got[z0:z0 + 64].astype(np.float64) - f * u0[z0:z0 + 64].astype(np.float64))
            assert float(d.max()) <= 1e-6, (steps, z0, float(d.max()))
        # the step changed the field by much more than the tolerance
        assert abs(1.0 - f) * float(np.abs(u0).max()) > 100e-6

    cnv.diffusion_step(a, b, dt)
    check(b, 1)
    res = cnv.diffusion_run(a, b, 2, dt)
    check(res, 2)
