"""CPU pins of the oracle's stencil operators (no GPU).

The 7-point Laplacian is pinned by the exact eigenvalue of a discrete
Fourier mode (a dropped term, sign or wrong neighbour fails it) and by an
independent whole-array evaluation with periodic rolls in the operation order
of Laplacian.cpp:125-129; the Euler update by its two separate roundings
(Field.h:457-512); the FP64 sum of squares by math.fsum.  WENO5 is pinned by
polynomial exactness, mirror symmetry, the linear weights on smooth data,
fifth-order convergence and ENO behaviour at a jump (Jiang & Shu 1996).
"""
import math

import numpy as np
import pytest

import oracle
from cubismnova_amd.layout import blocks_to_global, global_to_blocks
from tests.helpers import random_slab


def periodic_grid(nb, n, u, dtype=np.float64):
    return oracle.Field(global_to_blocks(u.astype(dtype), nb, n), nb, n)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 2e-3)])
def test_laplacian_fourier_eigenvalue(dtype, tol):
    nb, n = (2, 3, 2), (8, 4, 8)
    N = [nb[d] * n[d] for d in range(3)]
    h = [1.0 / N[d] for d in range(3)]
    k = (1, 2, 3)
    x = [(np.arange(N[d]) + 0.5) * h[d] for d in range(3)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    u = np.cos(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) + 0.3)
    lam = sum((2 * math.cos(2 * math.pi * k[d] * h[d]) - 2) / h[d] ** 2 for d in range(3))
    lap = blocks_to_global(oracle.laplacian(periodic_grid(nb, n, u, dtype), h), nb, n)
    err = np.max(np.abs(lap.astype(np.float64) - lam * u)) / abs(lam)
    assert err < tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_laplacian_roll_order_bitexact(dtype):
    nb, n = (3, 2, 2), (6, 8, 4)
    N = [nb[d] * n[d] for d in range(3)]
    h = [1.0 / N[d] for d in range(3)]
    u = random_slab(3, nb, n, dtype)
    G = blocks_to_global(u, nb, n)
    ih2 = [dtype(1.0 / (hh * hh)) for hh in h]
    two = dtype(2)
    ax = {0: 2, 1: 1, 2: 0}
    dd = []
    for d in range(3):
        p = np.roll(G, -1, axis=ax[d])
        m = np.roll(G, 1, axis=ax[d])
        dd.append(ih2[d] * ((p - two * G) + m))
    want = (dd[0] + dd[1]) + dd[2]
    got = blocks_to_global(oracle.laplacian(oracle.Field(u, nb, n), h), nb, n)
    assert got.dtype == dtype
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_euler_two_roundings_and_sumsq(dtype):
    nb, n = (2, 2, 2), (8, 8, 8)
    h = [1.0 / 16] * 3
    u = random_slab(9, nb, n, dtype)
    nu_dt = 0.1 * h[0] ** 2
    f = oracle.Field(u, nb, n)
    un, s, lap = oracle.diffusion_step(f, h, nu_dt, want_lap=True)
    assert np.array_equal(lap, oracle.laplacian(f, h))
    c = dtype(nu_dt)
    want = u + (lap * c)
    assert np.array_equal(un, want)
    ref = math.fsum(float(v) * float(v) for v in un.astype(np.float64))
    assert abs(s - ref) <= 1e-15 * ref


def test_diffusion_threads_deterministic():
    nb, n = (4, 4, 4), (8, 8, 8)
    u = random_slab(4, nb, n)
    f = oracle.Field(u, nb, n)
    a, sa = oracle.diffusion_step(f, [1 / 32] * 3, 1e-5, nthreads=1)
    b, sb = oracle.diffusion_step(f, [1 / 32] * 3, 1e-5, nthreads=4)
    assert np.array_equal(a, b) and sa == sb


def test_periodic_diffusion_conserves_sum():
    nb, n = (2, 2, 2), (8, 8, 8)
    u = random_slab(8, nb, n)
    f = oracle.Field(u, nb, n)
    un, _ = oracle.diffusion_step(f, [1 / 16] * 3, 0.1 / 256)
    assert abs(un.sum() - u.sum()) < 1e-12 * np.abs(u).sum()


# ------------------------------------------------------------------ WENO5


def cell_avg_poly(c, j):
    # average of c0 + c1 x + c2 x^2 over [j - 1/2, j + 1/2]
    return c[0] + c[1] * j + c[2] * (j * j + 1.0 / 12.0)


def poly(c, x):
    return c[0] + c[1] * x + c[2] * x * x


@pytest.mark.parametrize("coef", [(1.0, 0.0, 0.0), (0.3, -2.0, 0.0), (1.5, 0.7, -0.4), (-2.0, 3.0, 5.0)])
@pytest.mark.parametrize("tb", [False, True], ids=["diff", "textbook"])
def test_weno5_quadratic_exact(coef, tb):
    j = np.arange(-2, 3, dtype=np.float64)
    v = cell_avg_poly(coef, j)
    um, up = oracle.weno5_point(v, tb)
    assert abs(um - poly(coef, 0.5)) < 1e-12 * (1 + abs(poly(coef, 0.5)))
    assert abs(up - poly(coef, -0.5)) < 1e-12 * (1 + abs(poly(coef, -0.5)))


@pytest.mark.parametrize("tb", [False, True], ids=["diff", "textbook"])
def test_weno5_mirror_symmetry(tb):
    rng = np.random.Generator(np.random.PCG64(1))
    for _ in range(200):
        v = rng.standard_normal(5)
        um, up = oracle.weno5_point(v, tb)
        um2, up2 = oracle.weno5_point(v[::-1].copy(), tb)
        assert abs(um - up2) <= 1e-14 * (1 + abs(um)) and abs(up - um2) <= 1e-14 * (1 + abs(up))


@pytest.mark.parametrize("tb", [False, True], ids=["diff", "textbook"])
def test_weno5_linear_weights_on_smooth_data(tb):
    # smooth data: the result approaches the optimal 5-point linear scheme
    h = 1e-3
    j = np.arange(-2, 3)
    v = np.array([(np.sin((k + 0.5) * h) - np.sin((k - 0.5) * h)) / h for k in j])
    a, b, c, d, e = v
    lin = (2 * a - 13 * b + 47 * c + 27 * d - 3 * e) / 60.0
    um, _ = oracle.weno5_point(v, tb)
    assert abs(um - lin) < 1e-10


@pytest.mark.parametrize("tb", [False, True], ids=["diff", "textbook"])
def test_weno5_fifth_order(tb):
    errs = []
    for N in (32, 64, 128):
        h = 2 * np.pi / N
        xs = np.arange(N) * h
        avg = (np.cos(xs - h / 2) - np.cos(xs + h / 2)) / h  # cell averages of sin
        err = 0.0
        for i in range(N):
            v = avg[[(i + k) % N for k in range(-2, 3)]]
            um, _ = oracle.weno5_point(v, tb)
            err = max(err, abs(um - np.sin(xs[i] + h / 2)))
        errs.append(err)
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert orders[-1] > 4.5, (errs, orders)


@pytest.mark.parametrize("tb", [False, True], ids=["diff", "textbook"])
def test_weno5_eno_at_jump(tb):
    um, up = oracle.weno5_point(np.array([0.0, 0.0, 0.0, 1.0, 1.0]), tb)
    assert abs(um) < 1e-6          # stencil {i-2,i-1,i} dominates: no overshoot
    um, up = oracle.weno5_point(np.array([0.0, 0.0, 1.0, 1.0, 1.0]), tb)
    assert abs(up - 1.0) < 1e-6     # right state taken from the smooth side


def test_weno5_two_statements_agree():
    """The difference form (what the kernels evaluate) and the textbook
    statement give the same states to rounding, 1e-14 of the stencil's
    largest magnitude, over smooth, rough, discontinuous, constant and zero
    stencils at magnitudes from 1e-30 to 1e30."""
    rng = np.random.Generator(np.random.PCG64(7))
    cases = [np.zeros(5), np.full(5, 3.25), np.array([0.0, 0.0, 0.0, 1.0, 1.0]), np.array([1.0, 1.0, 0.0, 0.0, 0.0])]
    for _ in range(3000):
        kind = rng.integers(3)
        if kind == 0:
            v = rng.standard_normal(5)
        elif kind == 1:
            v = np.sin(rng.uniform(0, 6) + np.arange(5) * rng.uniform(1e-3, 0.5)) + rng.uniform(-2, 2)
        else:
            v = np.where(np.arange(5) < rng.integers(1, 5), rng.uniform(-1, 1), rng.uniform(-10, 10)) + 0.0
        cases.append(v * 10.0 ** rng.uniform(-30, 30))
    for v in cases:
        a = oracle.weno5_point(v)
        b = oracle.weno5_point(v, textbook=True)
        tol = 1e-14 * max(np.max(np.abs(v)), 1e-300)
        assert abs(a[0] - b[0]) <= tol and abs(a[1] - b[1]) <= tol, (v, a, b)


def test_weno5_field_matches_points():
    nb, n = (2, 1, 1), (6, 4, 4)
    u = random_slab(2, nb, n)
    f = oracle.Field(u, nb, n)
    outs = oracle.weno5(f, prec=1)
    G = blocks_to_global(u, nb, n)
    O = [blocks_to_global(o, nb, n) for o in outs]
    z, y, x = 2, 1, 11
    vx = np.array([G[z, y, (x + k) % 12] for k in range(-2, 3)])
    vy = np.array([G[z, (y + k) % 4, x] for k in range(-2, 3)])
    vz = np.array([G[(z + k) % 4, y, x] for k in range(-2, 3)])
    for i, v in enumerate((vx, vy, vz)):
        um, up = oracle.weno5_point(v)
        assert O[2 * i][z, y, x] == um and O[2 * i + 1][z, y, x] == up


# Known answers for the smoothness indicators, worked by hand from Jiang &
# Shu (1996), eq. (3.2)-(3.4): beta_k = 13/12 t_k^2 + 1/4 s_k^2 with the
# second differences t_k and the one-sided first differences s_k of the
# three-cell stencils of (a, b, c, d, e) = cells i-2 .. i+2.
#  * cell averages of x^3 on unit cells j = -2..2 (j^3 + j/4):
#    (-17/2, -5/4, 0, 5/4, 17/2); t = (-6, 0, 6), s = (-7/2, -5/2, 7/2);
#    beta = (36*13/12 + 49/16, 25/16, 36*13/12 + 49/16) = (673/16, 25/16, 673/16)
#  * cell averages of (x+1)^3: (-5/4, 0, 5/4, 17/2, 111/4);
#    t = (0, 6, 12), s = (5/2, -17/2, -5/2); beta = (25/16, 913/16, 2521/16)
#  * a kink (0, 1, 3, 2, 0): t = (1, -3, -1), s = (5, -1, 1);
#    beta = (13/12 + 25/4, 117/12 + 1/4, 13/12 + 1/4) = (22/3, 10, 4/3)
from fractions import Fraction as Fr

W5_KNOWN = [
    ((-8.5, -1.25, 0.0, 1.25, 8.5), (Fr(673, 16), Fr(25, 16), Fr(673, 16))),
    ((-1.25, 0.0, 1.25, 8.5, 27.75), (Fr(25, 16), Fr(913, 16), Fr(2521, 16))),
    ((0.0, 1.0, 3.0, 2.0, 0.0), (Fr(22, 3), Fr(10), Fr(4, 3))),
]


def w5_states_from_betas(v, beta):
    """Left/right states from given indicators, exactly: alpha_k =
    d_k / (eps + beta_k)^2 with eps = 1e-6 * m2 (m2 = mean of the five
    squares + 1e-30), normalised weights, and the third-order candidates of
    Jiang & Shu eq. (2.11) / their mirror images."""
    a, b, c, d, e = (Fr(x) for x in v)
    m2 = (a * a + b * b + c * c + d * d + e * e) / 5 + Fr(1e-30)
    eps = Fr(1e-6) * m2
    qm = ((2 * a - 7 * b + 11 * c) / 6, (-b + 5 * c + 2 * d) / 6, (2 * c + 5 * d - e) / 6)
    qp = ((-a + 5 * b + 2 * c) / 6, (2 * b + 5 * c - d) / 6, (11 * c - 7 * d + 2 * e) / 6)

    def combine(dk, q):
        al = [dk[k] / (eps + beta[k]) ** 2 for k in range(3)]
        return sum(al[k] * q[k] for k in range(3)) / sum(al)

    return (combine((Fr(1, 10), Fr(6, 10), Fr(3, 10)), qm), combine((Fr(3, 10), Fr(6, 10), Fr(1, 10)), qp))


def w5_betas_with(v, k13, k14):
    a, b, c, d, e = (Fr(x) for x in v)
    t = (a - 2 * b + c, b - 2 * c + d, c - 2 * d + e)
    s = (a - 4 * b + 3 * c, b - d, 3 * c - 4 * d + e)
    return tuple(k13 * t[k] ** 2 + k14 * s[k] ** 2 for k in range(3))


@pytest.mark.parametrize("case", range(len(W5_KNOWN)))
@pytest.mark.parametrize("tb", [False, True], ids=["diff", "textbook"])
def test_weno5_known_indicators(case, tb):
    v, beta = W5_KNOWN[case]
    um, up = w5_states_from_betas(v, beta)
    got = oracle.weno5_point(np.array(v), tb)
    for g, w in zip(got, (um, up)):
        assert abs(Fr(g) - w) <= Fr(1e-14) * (1 + abs(w)), (v, got, float(um), float(up))


@pytest.mark.parametrize("case", range(len(W5_KNOWN)))
def test_weno5_known_answers_see_the_constants(case):
    """The known answers above move by far more than their tolerance when
    either indicator constant is changed (13/12 -> 13/11, 1/4 -> 1/3), so a
    wrong constant shared by both oracle statements fails them."""
    v, beta = W5_KNOWN[case]
    assert w5_betas_with(v, Fr(13, 12), Fr(1, 4)) == beta
    ref = w5_states_from_betas(v, beta)
    for k13, k14 in ((Fr(13, 11), Fr(1, 4)), (Fr(13, 12), Fr(1, 3))):
        bad = w5_states_from_betas(v, w5_betas_with(v, k13, k14))
        moved = max(abs(bad[i] - ref[i]) / (1 + abs(ref[i])) for i in range(2))
        assert moved > Fr(1e-8), (v, k13, k14, float(moved))
