"""GPU parity of the batched launches: every grid of a batch of uneven grids
gives the result of its single-grid entry point (diffusion bit-exact, also
against the oracle; WENO5, whose FP32 evaluation contracts multiply-adds and
uses the hardware reciprocal, within 1e-6 of the single-grid kernel's largest
state and within 1e-6 rel-L2 of the oracle's FP64 evaluation), whether its
items run in an (operator, edge) class launch or in the persistent launch,
and a graph replay of the batch reproduces the eager run every time."""
import numpy as np
import pytest

import cubismnova_amd as cnv
import oracle
from cubismnova_amd import workloads
from tests.helpers import cells_of

pytestmark = pytest.mark.gpu


def make_grids(spec):
    nb, e = spec["nblocks"], spec["edge"]
    g = cnv.Cartesian(nb, e, dtype=np.float32, ncomp=spec["ncomp"])
    for c in range(spec["ncomp"]):
        g.upload_global(c, workloads.batch_field(spec, c).astype(np.float32))
    for kind, d, s, v in spec["bcs"]:
        g.addFaceBC(cnv.Dirichlet(d, s, v) if kind == "dirichlet" else cnv.Absorbing(d, s))
    nout = 6 * spec["ncomp"] if spec["op"] == "weno5" else spec["ncomp"]
    return g, cnv.Cartesian(nb, e, dtype=np.float32, ncomp=nout)


def oracle_bcs(spec):
    nb = spec["nblocks"]
    lists = {}
    for kind, d, s, v in spec["bcs"]:
        for b in range(nb[0] * nb[1] * nb[2]):
            bi = (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
            if bi[d] == (0 if s == 0 else nb[d] - 1):
                lists.setdefault(b, []).append((2 if kind == "dirichlet" else 1, d, s, v))
    return lists


@pytest.fixture(scope="module")
def batch_setup():
    specs = workloads.batch_specs(ngrids=14, k=40, max_blocks=(3, 3, 2), edges=(8, 16, 32))
    grids = [make_grids(s) for s in specs]
    items = []
    for s, (gi, go) in zip(specs, grids):
        h = 1.0 / (s["nblocks"][0] * s["edge"])
        items.append((gi, go, s["op"], s["ghost"], 0.1 * h * h))
    return specs, grids, items


def test_batch_matches_single_grid_and_oracle(gpu, batch_setup):
    specs, grids, items = batch_setup
    b = cnv.Batch(items)
    b.run()
    for s, (gi, go), it in zip(specs, grids, items):
        nb, e = s["nblocks"], s["edge"]
        ref = cnv.Cartesian(nb, e, dtype=np.float32, ncomp=go.ncomp)
        if s["op"] == "diffusion":
            cnv.diffusion_step(gi, ref, it[4])
        else:
            cnv.weno5(gi, ref)
        bcs = oracle_bcs(s)
        for c in range(s["ncomp"]):
            slab = gi.download(c)
            f = oracle.Field(slab, nb, e, gi.block_stride, bcs)
            if s["op"] == "diffusion":
                got = cells_of(go.download(c), nb, (e,) * 3, gi.block_stride)
                assert np.array_equal(got, cells_of(ref.download(c), nb, (e,) * 3, gi.block_stride))
                hh = [1.0 / (nb[d] * e) for d in range(3)]
                un, _ = oracle.diffusion_step(f, hh, it[4])
                assert np.array_equal(got, cells_of(un, nb, (e,) * 3, gi.block_stride))
            else:
                o64 = oracle.weno5(f, prec=1)
                for k in range(6):
                    got = cells_of(go.download(6 * c + k), nb, (e,) * 3, gi.block_stride).astype(np.float64)
                    one = cells_of(ref.download(6 * c + k), nb, (e,) * 3, gi.block_stride).astype(np.float64)
                    want = cells_of(o64[k], nb, (e,) * 3, gi.block_stride)
                    scale = np.max(np.abs(want))
                    assert np.max(np.abs(got - one)) <= 1e-6 * scale, (s["edge"], c, k)
                    assert np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want)


def test_batch_classes(gpu, batch_setup):
    specs, grids, items = batch_setup
    b = cnv.Batch(items)
    cls = b.classes()
    want = sorted({(0 if s["op"] == "diffusion" else 1, s["edge"]) for s in specs if s["edge"] in (16, 32)})
    assert sorted((c[0], c[1]) for c in cls) == want
    n_class = sum(c[4] for c in cls)
    n8 = sum(s["nblocks"][0] * s["nblocks"][1] * s["nblocks"][2] * s["ncomp"] for s in specs if s["edge"] == 8)
    assert b.info()["items"] >= n_class + n8      # 8^3 blocks: at least one item each
    for op, edge, wg, z, n in cls:
        assert wg in (64, 128, 256) and 1 <= z <= edge and n > 0


def test_batch_autotune_and_graph_replays(gpu, batch_setup):
    specs, grids, items = batch_setup
    b = cnv.Batch(items)
    b.run()
    builtin = [[go.download(c) for c in range(go.ncomp)] for _, go in grids]
    table = b.autotune(iters=2)
    assert len(table) > 0
    b.run()
    eager = [[go.download(c) for c in range(go.ncomp)] for _, go in grids]
    # the tuned class and persistent launches give the built-in choices' bits
    for ref, got in zip(builtin, eager):
        for c in range(len(ref)):
            assert np.array_equal(got[c], ref[c]), c
    st = cnv.Stream()
    gr = cnv.Graph(st).capture()
    b.run(stream=st)
    gr.end()
    for rep in range(3):
        for _, go in grids:            # clobber the outputs between replays
            for c in range(go.ncomp):
                cnv.call("cnv_fill", go.data_ptr(c), 0, -7.0, go.component_bytes // 4, None)
        gr.launch()
        st.synchronize()
        for (_, go), ref in zip(grids, eager):
            for c in range(go.ncomp):
                assert np.array_equal(go.download(c), ref[c]), (rep, c)


def test_batch_rejects_fp64_and_small_ghost(gpu):
    g = cnv.Cartesian(1, 8, dtype=np.float64)
    o = cnv.Cartesian(1, 8, dtype=np.float64)
    with pytest.raises(cnv.CnvError):
        cnv.Batch([(g, o, "diffusion", 1, 1e-3)])
    g32 = cnv.Cartesian(1, 8, dtype=np.float32)
    o32 = cnv.Cartesian(1, 8, dtype=np.float32, ncomp=6)
    with pytest.raises(cnv.CnvError):
        cnv.Batch([(g32, o32, "weno5", 1, 0.0)])


def test_full_batch_96_grids(gpu):
    """The full 96-grid workload (blocks 1x1x1 .. 8x8x4, edges {8,16,32},
    1-3 components, ghost 1 or 3, per-face periodic / Dirichlet / absorbing):
    after autotuning, a graph replay of the batch gives every grid's
    single-grid result (diffusion bit-exact, also against the oracle; WENO5
    within 1e-6 of the single-grid kernel's largest state and within 1e-6
    rel-L2 of the oracle's FP64 states), on each of three replays."""
    specs = workloads.batch_specs(96, k=4)
    grids = [make_grids(s) for s in specs]
    items = []
    for s, (gi, go) in zip(specs, grids):
        h = 1.0 / (s["nblocks"][0] * s["edge"])
        items.append((gi, go, s["op"], s["ghost"], 0.1 * h * h))
    b = cnv.Batch(items)
    assert b.info()["items"] > 0
    b.autotune(iters=2)
    st = cnv.Stream()
    b.run(stream=st)
    st.synchronize()
    gr = cnv.Graph(st).capture()
    b.run(stream=st)
    gr.end()
    nt = oracle.max_threads()
    want = []
    for s, (gi, go), it in zip(specs, grids, items):
        nb, e = s["nblocks"], s["edge"]
        ref = cnv.Cartesian(nb, e, dtype=np.float32, ncomp=go.ncomp)
        if s["op"] == "diffusion":
            cnv.diffusion_step(gi, ref, it[4])
        else:
            cnv.weno5(gi, ref)
        bcs = oracle_bcs(s)
        per = []
        for c in range(s["ncomp"]):
            f = oracle.Field(gi.download(c), nb, e, gi.block_stride, bcs)
            if s["op"] == "diffusion":
                one = cells_of(ref.download(c), nb, (e,) * 3, gi.block_stride)
                un, _ = oracle.diffusion_step(f, [1.0 / (nb[d] * e) for d in range(3)], it[4], nthreads=nt)
                assert np.array_equal(one, cells_of(un, nb, (e,) * 3, gi.block_stride)), (s["edge"], nb, c)
                per.append((one, None))
            else:
                o64 = oracle.weno5(f, prec=1, nthreads=nt)
                for k in range(6):
                    one = cells_of(ref.download(6 * c + k), nb, (e,) * 3, gi.block_stride)
                    per.append((one, cells_of(o64[k], nb, (e,) * 3, gi.block_stride)))
        want.append(per)
        del ref
    for rep in range(3):
        for _, go in grids:            # clobber the outputs between replays
            for c in range(go.ncomp):
                cnv.call("cnv_fill", go.data_ptr(c), 0, -7.0, go.component_bytes // 4, None)
        gr.launch()
        st.synchronize()
        for s, (gi, go), per in zip(specs, grids, want):
            nb, e = s["nblocks"], s["edge"]
            for k, (one, o64) in enumerate(per):
                got = cells_of(go.download(k), nb, (e,) * 3, gi.block_stride)
                if o64 is None:
                    assert np.array_equal(got, one), (rep, s["edge"], nb, k)
                else:
                    g64, one64 = got.astype(np.float64), one.astype(np.float64)
                    scale = np.max(np.abs(o64))
                    assert np.max(np.abs(g64 - one64)) <= 1e-6 * scale, (rep, s["edge"], nb, k)
                    assert np.linalg.norm(g64 - o64) <= 1e-6 * np.linalg.norm(o64), (rep, s["edge"], nb, k)


def test_batch_single_launch_form(gpu, batch_setup):
    """Before tuning a batch runs as one persistent launch over its work
    units, with every grid's single-grid result; the autotuner times both
    forms and the one it keeps gives the same bits."""
    specs, grids, items = batch_setup
    b = cnv.Batch(items)
    one, nunits = b.form()
    assert one == 1 and nunits > 0
    b.run()
    first = [[go.download(c) for c in range(go.ncomp)] for _, go in grids]
    for s, (gi, go), it in zip(specs, grids, items):
        ref = cnv.Cartesian(s["nblocks"], s["edge"], dtype=np.float32, ncomp=go.ncomp)
        if s["op"] == "diffusion":
            cnv.diffusion_step(gi, ref, it[4])
        else:
            cnv.weno5(gi, ref)
        n = (s["edge"],) * 3
        for c in range(go.ncomp):
            got = cells_of(go.download(c), s["nblocks"], n, gi.block_stride)
            want = cells_of(ref.download(c), s["nblocks"], n, gi.block_stride)
            if s["op"] == "diffusion":
                assert np.array_equal(got, want), (s["edge"], c)
            else:
                assert np.max(np.abs(got.astype(np.float64) - want)) <= 1e-6 * np.max(np.abs(want)), (s["edge"], c)
    table = b.autotune(iters=2)
    forms = {int(t[2]): t[4] for t in table if t[0] == -3}
    assert set(forms) == {0, 1, 2} and all(v > 0 for v in forms.values())
    assert b.form()[0] == min(forms, key=lambda k: (forms[k], k != 1))
    b.run()
    for (_, go), ref in zip(grids, first):
        for c in range(go.ncomp):
            assert np.array_equal(go.download(c), ref[c]), c
    # every form, pinned, gives the same bits
    for f in (cnv.Batch.CLASSES, cnv.Batch.CLASSES_SERIAL, cnv.Batch.SINGLE):
        b.set_form(f)
        assert b.form()[0] == f
        b.run()
        for (_, go), ref in zip(grids, first):
            for c in range(go.ncomp):
                assert np.array_equal(go.download(c), ref[c]), (f, c)
    with pytest.raises(cnv.CnvError):
        b.set_form(3)
