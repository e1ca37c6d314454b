"""The oracle against the committed known-answer fixtures: labs on the cells
the reference defines, and the Laplacian / Euler outputs, bit-exact in FP64
and FP32.  The fixtures come from an independent cell-by-cell evaluator
(tests/golden/make_fixtures.py)."""
import numpy as np
import pytest

import oracle
from tests import golden_data
from tests.helpers import defined_mask

Z, META = golden_data.load()


def rounded_stride(cells, dtype):
    n = int(np.prod(cells))
    isz = np.dtype(dtype).itemsize
    return ((n * isz + 31) // 32 * 32) // isz


def low_corner(a, shape, fill):
    """A block's lab placed in the low corner of the largest block's lab
    geometry (node / face blocks below the upper end are one smaller)."""
    if a.shape == tuple(shape):
        return a
    out = np.full(shape, fill, dtype=a.dtype)
    out[:a.shape[0], :a.shape[1], :a.shape[2]] = a
    return out


@pytest.mark.parametrize("entry", META, ids=[m["name"] for m in META])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_labs_match_fixture(entry, dtype):
    nb, n = entry["nblocks"], entry["cells"]
    el = golden_data.elements(entry)
    bs = rounded_stride(el, dtype)
    slab = golden_data.padded(Z[entry["name"] + "/input"], nb, el, bs, dtype)
    f = oracle.Field(slab, nb, n, bs, golden_data.bc_lists(entry), plus=golden_data.plus(entry))
    for li, L in enumerate(entry["labs"]):
        want = Z["%s/lab%d" % (entry["name"], li)]
        for b in range(want.shape[0]):
            bi = golden_data.block_index(b, nb)
            got = low_corner(oracle.lab_load(f, bi, L["begin"], L["end"], L["tensorial"]), want[b].shape, np.nan)
            m = ~np.isnan(want[b])
            assert np.array_equal(got[m], want[b][m].astype(dtype)), (entry["name"], li, bi)
            # the oracle's defined set contains the fixture's (it may define more:
            # values a BC copied from a cell no load has written stay undefined here)
            dm = low_corner(defined_mask(f, bi, L["begin"], L["end"], L["tensorial"]), want[b].shape, False)
            assert not np.any(m & ~dm), (entry["name"], li, bi)
            sb, e = L["begin"], golden_data.extent(entry, bi)
            assert m[-sb[2]:-sb[2] + e[2], -sb[1]:-sb[1] + e[1], -sb[0]:-sb[0] + e[0]].all()


DIFF = [m for m in META if "h" in m]


@pytest.mark.parametrize("entry", DIFF, ids=[m["name"] for m in DIFF])
@pytest.mark.parametrize("dtype,tag", [(np.float64, "f64"), (np.float32, "f32")])
def test_oracle_diffusion_matches_fixture(entry, dtype, tag):
    nb, n = entry["nblocks"], entry["cells"]
    bs = rounded_stride(n, dtype)
    slab = golden_data.padded(Z[entry["name"] + "/input"], nb, n, bs, dtype)
    f = oracle.Field(slab, nb, n, bs, golden_data.bc_lists(entry))
    nbt = int(np.prod(nb))
    lap = oracle.laplacian(f, entry["h"]).reshape(nbt, bs)[:, :int(np.prod(n))]
    un, _ = oracle.diffusion_step(f, entry["h"], entry["nu_dt"])
    un = un.reshape(nbt, bs)[:, :int(np.prod(n))]
    assert np.array_equal(lap.reshape(-1), Z["%s/lap_%s" % (entry["name"], tag)])
    assert np.array_equal(un.reshape(-1), Z["%s/euler_%s" % (entry["name"], tag)])
