"""FP64 Ewald sum of the 2D periodic gravity the torus solves: TEST
INFRASTRUCTURE ONLY (an independent reference for the oracle's and the HIP
path's P3M forces on the torus; the product path never touches it).

The force kernel of 2D gravity is f(d) = d / |d|^2 (d = source - target, in
physical units: x scaled by the aspect ratio, y by the height), the gradient
of ln|d|, with laplacian 2 pi delta. On a torus of area A the periodic solver
keeps no k = 0 mode (K/Scenario.kt:44-48: it subtracts a uniform background
holding all the mass), so the periodic kernel solves lap G = 2 pi (delta - 1/A):

    f_p(d) = sum over images n of exp(-alpha |d + n|^2) (d + n) / |d + n|^2
           + (2 pi / A) sum over k != 0 of k sin(k . d) exp(-k^2 / (4 alpha)) / k^2

(the short-range part e^{-alpha r^2} / r and its Fourier complement
(1 - e^{-alpha r^2}) / r, whose transform is 2 pi e^{-k^2/4alpha} / k). The
result does not depend on alpha, which the tests check.
"""
from __future__ import annotations

import numpy as np


def periodic_kernel(d: np.ndarray, lx: float, ly: float, alpha: float | None = None,
                    images: int = 4, kmax: int = 24) -> np.ndarray:
    """f_p at separations d (array (..., 2), physical units) on an lx x ly torus."""
    d = np.asarray(d, dtype=np.float64)
    if alpha is None:
        alpha = (3.0 / min(lx, ly)) ** 2
    out = np.zeros(d.shape, np.float64)
    for i in range(-images, images + 1):
        for j in range(-images, images + 1):
            e = d + np.array([i * lx, j * ly])
            r2 = (e * e).sum(-1)
            with np.errstate(invalid="ignore", divide="ignore"):
                w = np.where(r2 > 0, np.exp(-alpha * r2) / r2, 0.0)
            out += e * w[..., None]
    area = lx * ly
    for a in range(-kmax, kmax + 1):
        for b in range(-kmax, kmax + 1):
            if a == 0 and b == 0:
                continue
            k = np.array([2 * np.pi * a / lx, 2 * np.pi * b / ly])
            k2 = (k * k).sum()
            s = np.sin(d @ k) * np.exp(-k2 / (4 * alpha)) / k2
            out += (2 * np.pi / area) * s[..., None] * k
    return out


def accelerations(xy: np.ndarray, mass: np.ndarray, aspect: float, G: float, total_mass: float,
                  **kw) -> np.ndarray:
    """Accelerations in the reference's normalised units (per axis / viewport)
    of bodies at normalised positions xy on the torus: a_i = G / M sum_j m_j
    f_p(x_j - x_i), with x scaled by the aspect ratio."""
    p = np.asarray(xy, np.float64) * np.array([aspect, 1.0])
    d = p[None, :, :] - p[:, None, :]
    f = periodic_kernel(d, aspect, 1.0, **kw)
    idx = np.arange(len(p))
    f[idx, idx] = 0.0
    a = G / total_mass * (np.asarray(mass, np.float64)[None, :, None] * f).sum(1)
    return a / np.array([aspect, 1.0])
