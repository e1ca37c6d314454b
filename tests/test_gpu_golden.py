"""The HIP path against the committed known-answer fixtures (labs of every
block on the cells the reference defines; Laplacian and Euler step bit-exact
in FP64 and FP32)."""
import numpy as np
import pytest

import cubismnova_amd as cnv
from tests import golden_data

pytestmark = pytest.mark.gpu

Z, META = golden_data.load()
MK = {0: lambda d, s, v: cnv.Base(d, s), 1: lambda d, s, v: cnv.Absorbing(d, s),
      2: lambda d, s, v: cnv.Dirichlet(d, s, v), 3: lambda d, s, v: cnv.Symmetry(d, s, v)}


def entity_of(entry):
    """(entity, slab) holding the fixture's field: plus (1,1,1) is a node
    field, a single plus dimension d the faces normal to d (slab d)."""
    p = golden_data.plus(entry)
    if p == (0, 0, 0):
        return "cell", 0
    if p == (1, 1, 1):
        return "node", 0
    assert sum(p) == 1
    return "face", p.index(1)


def grid_for(entry, dtype):
    nb, n = entry["nblocks"], entry["cells"]
    ent, slab = entity_of(entry)
    g = cnv.Cartesian(nb, n, dtype=dtype, entity=ent)
    g.upload(slab, golden_data.padded(Z[entry["name"] + "/input"], nb, golden_data.elements(entry),
                                      g.block_stride, dtype))
    for b, lst in golden_data.bc_lists(entry).items():
        g.setBC(golden_data.block_index(b, nb), [MK[k](d, s, v) for (k, d, s, v) in lst], comp=slab)
    return g


CELL = [m for m in META if golden_data.plus(m) == (0, 0, 0)]
DIFF = [m for m in CELL if "h" in m]


@pytest.mark.parametrize("entry", META, ids=[m["name"] for m in META])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_labs_match_fixture(gpu, entry, dtype):
    g = grid_for(entry, dtype)
    _, slab = entity_of(entry)
    for li, L in enumerate(entry["labs"]):
        want = Z["%s/lab%d" % (entry["name"], li)]
        st = cnv.Stencil(L["begin"], L["end"], L["tensorial"])
        buf, ext = g.lab_all(st, comp=slab)
        buf.fill(dtype, np.nan)
        g.lab_all(st, comp=slab, out=buf)
        got = buf.download(dtype).reshape(g.nblocks_total, ext[2], ext[1], ext[0])
        assert got.shape == want.shape
        m = ~np.isnan(want)
        assert np.array_equal(got[m], want[m].astype(dtype)), (entry["name"], li)


@pytest.mark.parametrize("entry", DIFF, ids=[m["name"] for m in DIFF])
@pytest.mark.parametrize("dtype,tag", [(np.float64, "f64"), (np.float32, "f32")])
def test_diffusion_matches_fixture(gpu, entry, dtype, tag):
    g = grid_for(entry, dtype)
    nb, n = entry["nblocks"], entry["cells"]
    nbt, cells = int(np.prod(nb)), int(np.prod(n))
    lap = cnv.Cartesian(nb, n, dtype=dtype)
    cnv.laplacian(g, lap)
    out = cnv.Cartesian(nb, n, dtype=dtype)
    cnv.diffusion_step(g, out, entry["nu_dt"])
    lv = lap.download(0).reshape(nbt, g.block_stride)[:, :cells].reshape(-1)
    uv = out.download(0).reshape(nbt, g.block_stride)[:, :cells].reshape(-1)
    assert np.array_equal(lv, Z["%s/lap_%s" % (entry["name"], tag)])
    assert np.array_equal(uv, Z["%s/euler_%s" % (entry["name"], tag)])
