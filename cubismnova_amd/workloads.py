"""Seeded synthetic inputs for the stencil workloads (numpy, host side).

Every generator returns global (Z, Y, X) float64 arrays at cell centres
x = begin + (i + 1/2) h (Mesh/StructuredUniform.h:361-365); callers cast to
the grid dtype and upload with Cartesian.upload_global.  The seed of
workload k is 0x5EED0000 + k.
"""
import concurrent.futures
import math
import os

import numpy as np

SEED0 = 0x5EED0000


def rng(k, salt=0):
    return np.random.Generator(np.random.PCG64(SEED0 + k + (salt << 16)))


def cell_centres(ncells, begin=0.0, end=1.0):
    h = (end - begin) / ncells
    return begin + (np.arange(ncells) + 0.5) * h


def diffusion_small(N=64):
    """0.1*U[-1,1] noise + sin(2 pi x) cos(4 pi y) sin(6 pi z) on [0,1]^3."""
    r = rng(1)
    x = cell_centres(N)
    sx, cy, sz = np.sin(2 * np.pi * x), np.cos(4 * np.pi * x), np.sin(6 * np.pi * x)
    u = sz[:, None, None] * cy[None, :, None] * sx[None, None, :]
    return u + 0.1 * (2.0 * r.random((N, N, N)) - 1.0)


def fourier_modes(k, nmodes=8, kmax=16):
    """nmodes integer wavevectors with 1 <= |k| <= kmax, amplitude 1/|k|,
    uniform phases."""
    r = rng(k)
    modes = []
    while len(modes) < nmodes:
        kv = r.integers(-kmax, kmax + 1, size=3)
        kn = float(np.sqrt((kv * kv).sum()))
        if 1.0 <= kn <= kmax:
            modes.append((kv, 1.0 / kn, r.random() * 2 * np.pi))
    return modes


def fourier_field(N, modes, dtype=np.float64):
    """sum_m A_m cos(2 pi k_m . x + phi_m) on an N^3 grid, built plane by plane."""
    x = cell_centres(N)
    u = np.zeros((N, N, N), dtype=np.float64)
    for kv, amp, ph in modes:
        ax = np.exp(1j * 2 * np.pi * kv[0] * x)
        ay = np.exp(1j * 2 * np.pi * kv[1] * x)
        az = np.exp(1j * (2 * np.pi * kv[2] * x + ph))
        plane = ay[:, None] * ax[None, :]
        for iz in range(N):
            u[iz] += amp * (az[iz] * plane).real
    return u.astype(dtype, copy=False)


def diffusion_large(N=256):
    """Sum of 8 seeded Fourier modes, |k| in [1,16], amplitude 1/|k|."""
    return fourier_field(N, fourier_modes(2))


def diffusion_dt(h, nu=1.0):
    return 0.1 * h * h, nu


# ------------------------------------------------------------------ batch of grids


def batch_specs(ngrids=96, k=4, max_blocks=(8, 8, 4), edges=(8, 16, 32)):
    """Seeded descriptions of independent grids: block counts in
    [1, max_blocks], block edge from `edges`, 1-3 components, operator
    (diffusion step or WENO5; WENO5 uses ghost 3, diffusion ghost 1 or 3), and
    per axis either periodic or, per side, Dirichlet U[-1,1] / absorbing."""
    r = rng(k)
    specs = []
    for gi in range(ngrids):
        nb = tuple(int(r.integers(1, m + 1)) for m in max_blocks)
        edge = int(edges[r.integers(0, len(edges))])
        ncomp = int(r.integers(1, 4))
        op = "weno5" if r.random() < 0.5 else "diffusion"
        ghost = 3 if op == "weno5" else int((1, 3)[r.integers(0, 2)])
        bcs = []
        for d in range(3):
            if r.random() < 0.4:
                continue  # periodic axis
            for s in (0, 1):
                if r.random() < 0.5:
                    bcs.append(("dirichlet", d, s, float(r.uniform(-1, 1))))
                else:
                    bcs.append(("absorbing", d, s, 0.0))
        specs.append({"nblocks": nb, "edge": edge, "ncomp": ncomp, "op": op, "ghost": ghost, "bcs": bcs,
                      "seed": int(r.integers(0, 2 ** 31))})
    return specs


def batch_field(spec, comp):
    """Seeded Fourier-mode field (4 modes, |k| <= 4, amplitude 1/|k|) of one
    component of a batch grid, global (Z, Y, X) float64."""
    nb, e = spec["nblocks"], spec["edge"]
    shape = (nb[2] * e, nb[1] * e, nb[0] * e)
    r = np.random.Generator(np.random.PCG64(spec["seed"] + comp))
    xs = [cell_centres(n) for n in shape[::-1]]
    u = np.zeros(shape)
    for _ in range(4):
        kv = r.integers(-4, 5, size=3)
        kn = max(1.0, float(np.sqrt((kv * kv).sum())))
        ph = r.random() * 2 * np.pi
        u += np.cos(2 * np.pi * (kv[0] * xs[0][None, None, :] + kv[1] * xs[1][None, :, None]
                                 + kv[2] * xs[2][:, None, None]) + ph) / kn
    return u


# ------------------------------------------------------------------ Poisson source


def poisson_source(N, zc, k=5, nblobs=3, sigma=0.05):
    """nblobs seeded Gaussians (sigma, amplitudes U[-1,1], centres U[0.2,0.8]^3)
    on cell centres x, y uniform and z = zc (stretched), mean removed."""
    r = rng(k)
    x = cell_centres(N)
    f = np.zeros((N, N, N))
    for _ in range(nblobs):
        c = r.uniform(0.2, 0.8, 3)
        amp = r.uniform(-1, 1)
        gx = np.exp(-(x - c[0]) ** 2 / (2 * sigma ** 2))
        gy = np.exp(-(x - c[1]) ** 2 / (2 * sigma ** 2))
        gz = np.exp(-(np.asarray(zc) - c[2]) ** 2 / (2 * sigma ** 2))
        f += amp * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return f - f.mean()


# ------------------------------------------------------------------ WENO5 state

GAMMA = 1.4


def _sin_modes(r, nmodes, kmax):
    modes = []
    while len(modes) < nmodes:
        kv = r.integers(-kmax, kmax + 1, size=3)
        if (kv != 0).any():
            modes.append((kv, r.random() * 2 * np.pi))
    return modes


def weno_setup(k=3):
    """Seeded parameters of the compressible state: density modes, velocity
    modes per direction, and the unit normal of the pressure front."""
    r = rng(k)
    rho = _sin_modes(r, 4, 4)
    vel = [_sin_modes(r, 4, 4) for _ in range(3)]
    nrm = r.standard_normal(3)
    nrm /= np.linalg.norm(nrm)
    return rho, vel, nrm


def _modes_plane(modes, x, zc, amp):
    # amp * mean_m sin(2 pi k.x + phi) on the x-y plane at height zc, as
    # sin(a x) cos(b y + c) + cos(a x) sin(b y + c)
    out = np.zeros((x.size, x.size))
    for kv, ph in modes:
        ax = 2 * np.pi * kv[0] * x
        byc = 2 * np.pi * kv[1] * x + (2 * np.pi * kv[2] * zc + ph)
        out += np.outer(np.cos(byc), np.sin(ax))
        out += np.outer(np.sin(byc), np.cos(ax))
    return amp * out / len(modes)


def weno_component(N, comp, k=3):
    """Component comp of (rho, rho u, rho v, rho w, E) on an N^3 grid:
    rho = 1 + 0.2 sin-modes, velocity 0.1 random-phase modes, p = 1 + 9 s(x)
    with s = (1 + tanh(n.(x - 1/2) / (2h))) / 2 a planar front of width 2h
    (pressure ratio 10), E = p/(gamma-1) + rho |u|^2 / 2."""
    rho_m, vel_m, nrm = weno_setup(k)
    x = cell_centres(N)
    h = 1.0 / N
    out = np.empty((N, N, N), dtype=np.float64)

    def planes(iz0, iz1):
        for iz in range(iz0, iz1):
            zc = x[iz]
            rho = 1.0 + _modes_plane(rho_m, x, zc, 0.2)
            if comp == 0:
                out[iz] = rho
                continue
            vel = [_modes_plane(vel_m[d], x, zc, 0.1) for d in range(3)]
            if comp in (1, 2, 3):
                out[iz] = rho * vel[comp - 1]
                continue
            s = nrm[0] * (x[None, :] - 0.5) + nrm[1] * (x[:, None] - 0.5) + nrm[2] * (zc - 0.5)
            p = 1.0 + 9.0 * 0.5 * (1.0 + np.tanh(s / (2.0 * h)))
            out[iz] = p / (GAMMA - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)

    # planes are independent: chunks on threads (numpy releases the GIL in
    # the array arithmetic); every plane is computed the same way either way
    nth = max(1, min(16, len(os.sched_getaffinity(0)), N // 16))
    step = (N + nth - 1) // nth
    with concurrent.futures.ThreadPoolExecutor(max_workers=nth) as ex:
        list(ex.map(lambda i0: planes(i0, min(N, i0 + step)), range(0, N, step)))
    return out
