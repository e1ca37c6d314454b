"""CPU pins of the oracle's variable-coefficient 7-point operator and its
weighted-Jacobi sweep (no GPU).

The operator is checked against an independently assembled sparse matrix
(scipy) of the flux-form stencil with constant (Dirichlet 0) ghosts, against
the continuum Laplacian of a smooth function on interior cells (second-order
convergence on the tanh-stretched mesh), and for symmetry in the
volume-weighted inner product; the sweep by convergence to the direct
solution of that matrix."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from cubismnova_amd import stretch
from cubismnova_amd.layout import blocks_to_global, global_to_blocks
from tests.helpers import face_bc_lists

D0 = [(2, d, s, 0.0) for d in range(3) for s in (0, 1)]   # Dirichlet u = 0 on all faces


def field(u, nb, n, dtype=np.float64):
    return oracle.Field(global_to_blocks(u.astype(dtype), nb, n), nb, n, bcs=face_bc_lists(nb, D0))


def apply_A(u, nb, n, czm, czp, h):
    f = field(u, nb, n)
    r, _ = oracle.jacobi(f, np.zeros_like(f.slab), czm, czp, h, 2.0 / 3.0, 1)
    return -blocks_to_global(r, nb, n)


def matrix(N, czm, czp, h):
    """7-point matrix on an N^3 grid (x fastest), ghosts = 0."""
    ix2, iy2 = 1.0 / h[0] ** 2, 1.0 / h[1] ** 2
    idx = np.arange(N ** 3).reshape(N, N, N)
    rows, cols, vals = [], [], []
    diag = np.zeros((N, N, N))
    diag -= 2 * ix2 + 2 * iy2
    diag -= (czm + czp)[:, None, None]
    rows.append(idx.ravel()); cols.append(idx.ravel()); vals.append(diag.ravel())
    for ax, w in ((2, ix2), (1, iy2)):
        for s in (1, -1):
            src = np.roll(idx, -s, axis=ax)
            mask = np.ones((N, N, N), bool)
            sl = [slice(None)] * 3
            sl[ax] = -1 if s == 1 else 0
            mask[tuple(sl)] = False
            rows.append(idx[mask]); cols.append(src[mask]); vals.append(np.full(mask.sum(), w))
    for s, c in ((1, czp), (-1, czm)):
        src = np.roll(idx, -s, axis=0)
        mask = np.ones((N, N, N), bool)
        mask[-1 if s == 1 else 0] = False
        rows.append(idx[mask]); cols.append(src[mask])
        vals.append(np.broadcast_to(c[:, None, None], (N, N, N))[mask])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N ** 3,) * 2)


def test_operator_matches_sparse_matrix():
    N, nb, n = 16, (2, 2, 2), (8, 8, 8)
    czm, czp, _, _ = stretch.metric(stretch.tanh_faces(N, 2.0))
    h = [1.0 / N] * 3
    u = np.random.Generator(np.random.PCG64(3)).standard_normal((N, N, N))
    Au = apply_A(u, nb, n, czm, czp, h)
    want = (matrix(N, czm, czp, h) @ u.ravel()).reshape(N, N, N)
    assert np.max(np.abs(Au - want)) <= 1e-12 * np.max(np.abs(want))


def test_operator_symmetric_in_volume_weights():
    N, nb, n = 16, (2, 2, 2), (8, 8, 8)
    czm, czp, dz, _ = stretch.metric(stretch.tanh_faces(N, 2.0))
    h = [1.0 / N] * 3
    rng = np.random.Generator(np.random.PCG64(4))
    u, v = rng.standard_normal((2, N, N, N))
    w = dz[:, None, None]
    a = np.sum(w * v * apply_A(u, nb, n, czm, czp, h))
    b = np.sum(w * u * apply_A(v, nb, n, czm, czp, h))
    assert abs(a - b) <= 1e-12 * abs(a)


def test_operator_second_order_on_stretched_mesh():
    errs = []
    for N in (32, 64):
        nb, n = (N // 16,) * 3, (16, 16, 16)
        zf = stretch.tanh_faces(N, 2.0)
        czm, czp, _, zc = stretch.metric(zf)
        h = [1.0 / N] * 3
        x = (np.arange(N) + 0.5) / N
        k = np.pi
        u = np.sin(k * zc)[:, None, None] * np.sin(k * x)[None, :, None] * np.cos(k * x)[None, None, :]
        Au = apply_A(u, nb, n, czm, czp, h)
        s = slice(N // 8, N - N // 8)
        errs.append(np.max(np.abs(Au[s, s, s] + 3 * k * k * u[s, s, s])))
    order = np.log2(errs[0] / errs[1])
    assert order > 1.8, (errs, order)


def test_weighted_jacobi_converges_to_direct_solution():
    N, nb, n = 8, (1, 1, 1), (8, 8, 8)
    czm, czp, _, zc = stretch.metric(stretch.tanh_faces(N, 2.0))
    h = [1.0 / N] * 3
    rng = np.random.Generator(np.random.PCG64(5))
    f = rng.standard_normal((N, N, N))
    f -= f.mean()
    exact = spla.spsolve(matrix(N, czm, czp, h).tocsc(), f.ravel()).reshape(N, N, N)
    u = np.zeros((N, N, N))
    fs = global_to_blocks(f, nb, n)
    res = []
    for it in range(1000):
        if it % 250 == 0:
            r, s = oracle.jacobi(field(u, nb, n), fs, czm, czp, h, 2.0 / 3.0, 1)
            res.append(np.sqrt(s))
        u = blocks_to_global(oracle.jacobi(field(u, nb, n), fs, czm, czp, h, 2.0 / 3.0, 0), nb, n)
    # geometric decrease until the residual reaches rounding level
    assert res[1] < 1e-5 * res[0] and res[2] < 1e-5 * res[1]
    assert np.max(np.abs(u - exact)) <= 1e-12 * np.max(np.abs(exact))
