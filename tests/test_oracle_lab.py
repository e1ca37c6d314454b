"""CPU pins of the oracle's lab assembly (no GPU).

Each test reproduces a check of the reference's own unit tests
(test/UnitTests/Block/FieldLabTest.cpp, test/UnitTests/BC/*Test.cpp,
test/UnitTests/Grid/CartesianTest.cpp) or an independent formulation
(global-index periodic wrap, hand-derived known answers), so that a wrong
index, side or order in the oracle fails one of them.
"""
import itertools

import numpy as np
import pytest

import oracle
from tests.helpers import defined_mask, face_bc_lists, random_slab

B, A, D, S = oracle.BASE, oracle.ABSORBING, oracle.DIRICHLET, oracle.SYMMETRY


def lab_at(L, sb, q):
    return L[q[2] - sb[2], q[1] - sb[1], q[0] - sb[0]]


def iota_self_periodic(n=16, dtype=np.float64):
    # FieldAndLab / FieldLabTest fixture: one field filled 0,1,2,... that maps
    # every neighbour index to itself (a 1x1x1 periodic grid).
    slab = np.arange(n ** 3, dtype=np.float64).astype(dtype)
    return oracle.Field(slab, 1, n)


# ------------------------------------------------- FieldLabTest.cpp:16-115


@pytest.mark.parametrize("tensorial", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fieldlab_ghosts_periodic(tensorial, dtype):
    f = iota_self_periodic(16, dtype)
    sb, se = [-2] * 3, [3] * 3
    L = oracle.lab_load(f, (0, 0, 0), -2, 3, tensorial)
    n = 16
    field = f.slab.reshape(n, n, n)
    # inner data
    assert np.array_equal(L[2:2 + n, 2:2 + n, 2:2 + n], field)
    if not tensorial:
        for p in itertools.product(range(n), repeat=3):
            if not (0 in p or n - 1 in p):
                continue
            for i in range(3):
                e = np.eye(3, dtype=int)[i]
                if p[i] == 0:
                    for si in range(sb[i], 0):
                        assert lab_at(L, sb, p + si * e) == lab_at(L, sb, p + (si + n) * e)
                if p[i] == n - 1:
                    for si in range(1, se[i]):
                        assert lab_at(L, sb, p + (si - n) * e) == lab_at(L, sb, p + si * e)
    else:
        for q in itertools.product(range(-2, n + 2), repeat=3):
            if all(0 <= v < n for v in q):
                continue
            s = [v % n for v in q]
            assert lab_at(L, sb, q) == lab_at(L, sb, s)


# ------------------------------------------------- BC/*Test.cpp


def _slab_iter(extent, start, dir, depth=3):
    ext = list(extent)
    ext[dir] = depth
    for p in itertools.product(*[range(e) for e in ext[::-1]]):
        p = p[::-1]
        yield [p[d] + start[d] for d in range(3)]


@pytest.mark.parametrize("tensorial", [False, True])
def test_bc_absorbing(tensorial):
    f = iota_self_periodic()
    n = 16
    for dirs in ([0, 1, 2] if not tensorial else [[0], [1], [2]]):
        dl = dirs if isinstance(dirs, list) else [dirs]
        bcv = [(A, d, s, 0.0) for d in dl for s in (0, 1)]
        L = oracle.lab_load(f, (0, 0, 0), -3, 4, tensorial, extern=bcv)
        sb = [-3] * 3
        ext = [n + 6] * 3 if tensorial else [n] * 3
        base = sb if tensorial else [0, 0, 0]
        for d in dl:
            st = list(base)
            st[d] = -3
            for q in _slab_iter(ext, st, d):
                r = list(q)
                r[d] = 0
                assert lab_at(L, sb, q) == lab_at(L, sb, r)
            st[d] = n
            for q in _slab_iter(ext, st, d):
                r = list(q)
                r[d] = n - 1
                assert lab_at(L, sb, q) == lab_at(L, sb, r)


@pytest.mark.parametrize("tensorial", [False, True])
def test_bc_dirichlet(tensorial):
    f = iota_self_periodic()
    n = 16
    cases = [[(0, 10, 11), (1, 12, 13), (2, 14, 15)]] if not tensorial else \
        [[(0, 101, 102)], [(1, 103, 104)], [(2, 105, 106)]]
    for case in cases:
        bcv = []
        for d, v0, v1 in case:
            bcv += [(D, d, 0, v0), (D, d, 1, v1)]
        L = oracle.lab_load(f, (0, 0, 0), -3, 4, tensorial, extern=bcv)
        sb = [-3] * 3
        ext = [n + 6] * 3 if tensorial else [n] * 3
        base = sb if tensorial else [0, 0, 0]
        for d, v0, v1 in case:
            st = list(base)
            st[d] = -3
            assert all(lab_at(L, sb, q) == v0 for q in _slab_iter(ext, st, d))
            st[d] = n
            assert all(lab_at(L, sb, q) == v1 for q in _slab_iter(ext, st, d))


@pytest.mark.parametrize("tensorial", [False, True])
def test_bc_symmetry(tensorial):
    f = iota_self_periodic()
    n = 16
    cases = [[(0, 1, -1), (1, 1, -1), (2, 1, -1)]] if not tensorial else [[(0, 1, 1)], [(1, 1, 1)], [(2, 1, 1)]]
    for case in cases:
        bcv = []
        for d, s0, s1 in case:
            bcv += [(S, d, 0, s0), (S, d, 1, s1)]
        L = oracle.lab_load(f, (0, 0, 0), -3, 4, tensorial, extern=bcv)
        sb = [-3] * 3
        ext = [n + 6] * 3 if tensorial else [n] * 3
        base = sb if tensorial else [0, 0, 0]
        for d, s0, s1 in case:
            st = list(base)
            st[d] = -3
            for q in _slab_iter(ext, st, d):
                r = list(q)
                r[d] = -1 - q[d]
                assert lab_at(L, sb, q) == s0 * lab_at(L, sb, r)
            st[d] = n
            for q in _slab_iter(ext, st, d):
                r = list(q)
                r[d] = 2 * n - 1 - q[d]
                assert lab_at(L, sb, q) == s1 * lab_at(L, sb, r)


def test_bc_base_keeps_periodic():
    # BaseTest.cpp:16-67: a Base BC leaves the periodic ghosts
    f = iota_self_periodic()
    bcv = [(B, d, s, 0.0) for d in range(3) for s in (0, 1)]
    L1 = oracle.lab_load(f, (0, 0, 0), -3, 4, False, extern=bcv)
    L0 = oracle.lab_load(f, (0, 0, 0), -3, 4, False)
    m = defined_mask(f, (0, 0, 0), -3, 4)
    assert np.array_equal(L1[m], L0[m])


@pytest.mark.parametrize("kind,val", [(A, 0.0), (D, 1.0), (S, 1.0)])
def test_bc_no_stencil_width(kind, val):
    # CommonTest.cpp:20-61: Stencil(0,1) makes every BC a no-op
    slab = np.arange(64, dtype=np.float64)
    f = oracle.Field(slab, 1, 4)
    L = oracle.lab_load(f, (0, 0, 0), 0, 1, False, extern=[(kind, 0, 0, val)])
    assert np.array_equal(L, slab.reshape(4, 4, 4))


# ------------------------------------------------- CartesianTest.cpp:172-329


@pytest.mark.parametrize("dtype", [np.float64])
def test_cartesian_lab_analytic(dtype):
    nb, n = 3, 8
    N = nb * n
    h = 1.0 / N
    fexact = lambda x, y, z: np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)
    c = (np.arange(N) + 0.5) * h
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    from cubismnova_amd.layout import global_to_blocks
    slab = global_to_blocks(fexact(X, Y, Z), (nb,) * 3, (n,) * 3)
    f = oracle.Field(slab, nb, n)
    eps = np.finfo(np.float64).eps
    for bz, by, bx in itertools.product(range(nb), repeat=3):
        L = oracle.lab_load(f, (bx, by, bz), -2, 3, True)
        q = np.arange(-2, n + 2)
        xs = (bx * n + q + 0.5) * h
        ys = (by * n + q + 0.5) * h
        zs = (bz * n + q + 0.5) * h
        ZZ, YY, XX = np.meshgrid(zs, ys, xs, indexing="ij")
        assert np.max(np.abs(L - fexact(XX, YY, ZZ))) <= 8 * eps


def entity_slab(nb, n, plus, fexact, h):
    """Slab of a node / face field in the reference layout: block b stores its
    elements x-fastest in its own extent (n, plus one at the upper end of the
    plus dims), blocks (n+1)^3 elements apart (Cartesian.h:490-507,
    BlockFieldAssembler.h:121-131); element i of a plus dim sits at i h, of a
    cell dim at (i + 1/2) h."""
    stride = (n + 1) ** 3
    slab = np.zeros(nb ** 3 * stride)
    for bz, by, bx in itertools.product(range(nb), repeat=3):
        b = (bx, by, bz)
        ext = [n + (1 if plus[d] and b[d] == nb - 1 else 0) for d in range(3)]
        co = [(b[d] * n + np.arange(ext[d]) + (0.0 if plus[d] else 0.5)) * h for d in range(3)]
        ZZ, YY, XX = np.meshgrid(co[2], co[1], co[0], indexing="ij")
        flat = bx + nb * (by + nb * bz)
        slab[flat * stride:flat * stride + ext[0] * ext[1] * ext[2]] = fexact(XX, YY, ZZ).ravel()
    return slab, stride


@pytest.mark.parametrize("plus", [(1, 1, 1), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
                         ids=["node", "face_x", "face_y", "face_z"])
@pytest.mark.parametrize("tensorial", [True, False])
def test_node_face_lab_analytic(plus, tensorial):
    """CartesianTest.cpp:172-318 for node fields (and the same construction per
    face direction): 3^3 blocks of 8^3, Stencil(-2, 3), periodic.  Lab values
    equal the function at the element coordinates, where a coordinate of a
    plus dim outside [0, 1] is moved inwards by h -- the periodic images the
    shifted neighbour copies (FieldLabLoader.h:217-219) produce -- to 8 eps."""
    nb, n = 3, 8
    h = 1.0 / (nb * n)
    fx = lambda x, y, z: np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.sin(2 * np.pi * z)

    def fexact(X, Y, Z):
        c = [X, Y, Z]
        for d in range(3):
            if plus[d]:
                c[d] = np.where(c[d] < 0, c[d] + h, np.where(c[d] > 1, c[d] - h, c[d]))
        return fx(*c)

    slab, stride = entity_slab(nb, n, plus, fx, h)
    f = oracle.Field(slab, nb, n, stride, plus=plus)
    eps = np.finfo(np.float64).eps
    checked = 0
    for bz, by, bx in itertools.product(range(nb), repeat=3):
        b = (bx, by, bz)
        L = oracle.lab_load(f, b, -2, 3, tensorial)
        ext = oracle.block_extent(f, b)
        co = [(b[d] * n + np.arange(-2, ext[d] + 2) + (0.0 if plus[d] else 0.5)) * h for d in range(3)]
        ZZ, YY, XX = np.meshgrid(co[2], co[1], co[0], indexing="ij")
        m = defined_mask(f, b, -2, 3, tensorial)
        assert L.shape == XX.shape
        assert np.max(np.abs(L[m] - fexact(XX, YY, ZZ)[m])) <= 8 * eps
        checked += int(m.sum())
    assert checked > 0


# ------------------------------------------------- independent formulations


@pytest.mark.parametrize("tensorial", [False, True])
@pytest.mark.parametrize("sten", [(-1, 2), (-3, 4), ((-1, -2, -3), (3, 2, 1))])
def test_periodic_global_wrap(tensorial, sten):
    """lab of block B at local q equals the global array at (B*N + q) mod total."""
    nb, n = (3, 2, 4), (5, 4, 6)
    slab = random_slab(11, nb, n)
    f = oracle.Field(slab, nb, n)
    from cubismnova_amd.layout import blocks_to_global
    G = blocks_to_global(slab, nb, n)
    sb = [sten[0]] * 3 if np.isscalar(sten[0]) else list(sten[0])
    se = [sten[1]] * 3 if np.isscalar(sten[1]) else list(sten[1])
    tot = [nb[d] * n[d] for d in range(3)]
    for bz, by, bx in itertools.product(range(nb[2]), range(nb[1]), range(nb[0])):
        L = oracle.lab_load(f, (bx, by, bz), sb, se, tensorial)
        m = defined_mask(f, (bx, by, bz), sb, se, tensorial)
        qz = np.arange(sb[2], n[2] + se[2] - 1)
        qy = np.arange(sb[1], n[1] + se[1] - 1)
        qx = np.arange(sb[0], n[0] + se[0] - 1)
        gz = (bz * n[2] + qz) % tot[2]
        gy = (by * n[1] + qy) % tot[1]
        gx = (bx * n[0] + qx) % tot[0]
        want = G[np.ix_(gz, gy, gx)]
        # expected defined set: everything (tensorial) or inner + faces
        oz = ((qz < 0) | (qz >= n[2]))[:, None, None]
        oy = ((qy < 0) | (qy >= n[1]))[None, :, None]
        ox = ((qx < 0) | (qx >= n[0]))[None, None, :]
        nout = oz.astype(int) + oy.astype(int) + ox.astype(int)
        exp_mask = np.ones_like(m) if tensorial else (nout <= 1)
        assert np.array_equal(m, exp_mask)
        assert np.array_equal(L[m], want[m])


def test_face_bcs_closed_form():
    """Non-tensorial faces with one BC per face, derived per face by hand:
    x periodic, y-: Dirichlet 0.25, y+: absorbing, z-: symmetry -1, z+: Dirichlet -0.5."""
    nb, n = (2, 2, 2), (8, 8, 8)
    slab = random_slab(5, nb, n)
    from cubismnova_amd.layout import blocks_to_global
    G = blocks_to_global(slab, nb, n)
    bcs = face_bc_lists(nb, [(D, 1, 0, 0.25), (A, 1, 1, 0.0), (S, 2, 0, -1.0), (D, 2, 1, -0.5)])
    f = oracle.Field(slab, nb, n, bcs=bcs)
    g = 2
    for bz, by, bx in itertools.product(range(2), repeat=3):
        L = oracle.lab_load(f, (bx, by, bz), -g, g + 1, False)
        o = (bx * 8, by * 8, bz * 8)
        for q in itertools.product(range(-g, 8 + g), repeat=3):
            out = [v < 0 or v >= 8 for v in q]
            if sum(out) != 1:
                continue
            gq = [o[d] + q[d] for d in range(3)]
            if out[0]:  # periodic in x
                want = G[gq[2], gq[1], gq[0] % 16]
            elif out[1]:
                if by == 0 and q[1] < 0:
                    want = 0.25
                elif by == 1 and q[1] >= 8:
                    want = G[gq[2], 15, gq[0]]
                else:
                    want = G[gq[2], gq[1], gq[0]]
            else:
                if bz == 0 and q[2] < 0:
                    want = -1.0 * G[-1 - q[2], gq[1], gq[0]]
                elif bz == 1 and q[2] >= 8:
                    want = -0.5
                else:
                    want = G[gq[2], gq[1], gq[0]]
            assert lab_at(L, [-g] * 3, q) == want, (bx, by, bz, q)


def test_tensorial_corner_order_known_answer():
    """Tensorial edges are decided by the last BC in the list (FieldLab.h:219-223)."""
    slab = np.arange(8, dtype=np.float64) + 1.0  # one block of 2^3 cells: u(x,y,z)=1+x+2y+4z
    f = oracle.Field(slab, 1, 2)
    u = lambda x, y, z: 1.0 + (x % 2) + 2 * (y % 2) + 4 * (z % 2)
    # Dirichlet(x-)=5 then Absorbing(y-): the edge (-1,-1,z) copies (-1,0,z) = 5
    L = oracle.lab_load(f, (0, 0, 0), -1, 2, True, extern=[(D, 0, 0, 5.0), (A, 1, 0, 0.0)])
    for z in range(-1, 3):
        assert lab_at(L, [-1] * 3, (-1, -1, z)) == 5.0
        assert lab_at(L, [-1] * 3, (2, -1, z)) == u(2, 0, z)   # absorbing copies the periodic ghost
        assert lab_at(L, [-1] * 3, (1, -1, z)) == u(1, 0, z)
    # reversed order: Dirichlet wins on the edge, absorbing elsewhere
    L = oracle.lab_load(f, (0, 0, 0), -1, 2, True, extern=[(A, 1, 0, 0.0), (D, 0, 0, 5.0)])
    for z in range(-1, 3):
        assert lab_at(L, [-1] * 3, (-1, -1, z)) == 5.0
        assert lab_at(L, [-1] * 3, (0, -1, z)) == u(0, 0, z)
    # symmetry on x- then absorbing on y-: (-1,-1,z) = (-1,0,z) = sign*u(0,0,z)
    L = oracle.lab_load(f, (0, 0, 0), -1, 2, True, extern=[(S, 0, 0, -1.0), (A, 1, 0, 0.0)])
    for z in range(-1, 3):
        assert lab_at(L, [-1] * 3, (-1, -1, z)) == -u(0, 0, z)


def test_one_block_both_sides_known_answer():
    """1x1x1 grid with Dirichlet on both x sides: the skipped side is filled by
    its BC, the other side first loaded periodically then overwritten."""
    slab = np.arange(64, dtype=np.float64)
    bcs = {0: [(D, 0, 0, -1.0), (D, 0, 1, 2.0)]}
    f = oracle.Field(slab, 1, 4, bcs=bcs)
    L = oracle.lab_load(f, (0, 0, 0), -1, 2, False)
    F = slab.reshape(4, 4, 4)
    assert np.all(L[1:5, 1:5, 0] == -1.0)
    assert np.all(L[1:5, 1:5, 5] == 2.0)
    assert np.array_equal(L[1:5, 0, 1:5], F[:, 3, :])  # y periodic kept
    assert np.array_equal(L[0, 1:5, 1:5], F[3, :, :])


def test_unwritten_asymmetric_stencil():
    """Absorbing(x-) with a stencil of zero width on x+ is not applied
    (BC/Base.h:73-78) and its side is skipped: those ghosts stay unwritten."""
    slab = np.arange(27, dtype=np.float64)
    f = oracle.Field(slab, 1, 3, bcs={0: [(A, 0, 0, 0.0)]})
    m = defined_mask(f, (0, 0, 0), (-1, -1, -1), (1, 2, 2), False)
    assert not m[1:4, 1:4, 0].any()       # x = -1 face never written
    assert m[1:4, 1:4, 1:4].all()
    assert m[1:4, 0, 1:4].all() and m[0, 1:4, 1:4].all()
