"""Stretched mesh metric along z for the variable-coefficient 7-point operator.

CubismNova names a stretched mesh class (include/Cubism/Common.h:86,
MeshClass::Stretched) but ships no implementation; this module defines the
one the Jacobi solver uses.  Face positions follow a symmetric tanh map of
factor beta on [0, 1]:

    z_f(j) = (1 + tanh(beta (2 j / N - 1)) / tanh(beta)) / 2,  j = 0..N

Cell k spans [z_f(k), z_f(k+1)], its centre is the midpoint, dz_k the width.
The operator in z is the flux form

    (A u)_k = czp_k (u_{k+1} - u_k) - czm_k (u_k - u_{k-1}),
    czp_k = 1 / (dz_k * (zc_{k+1} - zc_k)),  czm_k = 1 / (dz_k * (zc_k - zc_{k-1})),

with the ghost centres outside the domain mirrored about the boundary face
(zc_{-1} = 2 z_f(0) - zc_0, zc_N = 2 z_f(N) - zc_{N-1}), which is where the
constant ghost value of a Dirichlet BC (BC/Dirichlet.h:75-113) sits.
"""
import numpy as np


def tanh_faces(N, beta=2.0):
    j = np.arange(N + 1, dtype=np.float64)
    return 0.5 * (1.0 + np.tanh(beta * (2.0 * j / N - 1.0)) / np.tanh(beta))


def uniform_faces(N):
    return np.arange(N + 1, dtype=np.float64) / N


def metric(zf):
    """(czm, czp, dz, zc) for face positions zf (length N+1)."""
    zf = np.asarray(zf, dtype=np.float64)
    dz = np.diff(zf)
    zc = 0.5 * (zf[1:] + zf[:-1])
    zg = np.concatenate(([2 * zf[0] - zc[0]], zc, [2 * zf[-1] - zc[-1]]))
    dcm = zg[1:-1] - zg[:-2]
    dcp = zg[2:] - zg[1:-1]
    return 1.0 / (dz * dcm), 1.0 / (dz * dcp), dz, zc
