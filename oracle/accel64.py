"""FP64 restatement of the P3M acceleration field and of its transposed
Jacobian: TEST INFRASTRUCTURE ONLY (the checker of the HIP adjoint).

A(x_i) = CIC gather of F = K * rho (rho the CIC deposit of the masses, K the
Gaussian-split force kernel of S/kernel_init.comp:16-46 with the
deconvolution of S/convolve.comp:24-47) + the near field of
S/kick.comp:66-80,118-151 (home cell: the first 8 list entries in
descending body index, cells with more than 2 bodies as monopoles at the
cell centre, the rest body by body), for a torus. vjp() is J_A(x)^T w with
the near-field topology frozen, by the decomposition the HIP adjoint uses;
the tests check it against finite differences of accel() in FP64.
Small N only: plain Python loops over bodies and stencil cells.
"""
from __future__ import annotations

import math

import numpy as np


class Accel64:
    def __init__(self, width, height, grid, gravity=0.035, softening=0.85, near_sigma=3.0,
                 near_min=1, near_max=6, exact_limit=2, home_limit=8, min_dist_frac=0.05,
                 window_floor=0.08):
        self.W, self.H = float(width), float(height)
        self.pw, self.ph = grid
        self.mw, self.mh = self.pw, self.ph
        self.asp = self.W / self.H
        self.G = gravity
        self.sigma = softening / self.ph
        self.k2 = 1.0 / (2.0 * self.sigma * self.sigma)
        cw, ch = self.asp / self.pw, 1.0 / self.ph
        md = min_dist_frac * min(cw, ch)
        self.md2 = md * md
        self.exact_limit, self.home_limit = exact_limit, home_limit
        # K/GravityPipeline.kt:153-169
        nx = near_sigma * softening * self.pw / (self.ph * self.asp)
        ny = softening * near_sigma
        self.rx = min(max(int(math.ceil(nx)), near_min), near_max)
        self.ry = min(max(int(math.ceil(ny)), near_min), near_max)
        self.window_floor = window_floor

    def _kernel(self, total_mass):
        mw, mh, pw, ph = self.mw, self.mh, self.pw, self.ph
        K = np.zeros((2, mh, mw))
        for ry in range(mh):
            for rx in range(mw):
                ox = rx if rx <= mw // 2 else rx - mw
                oy = ry if ry <= mh // 2 else ry - mh
                dx, dy = ox * self.asp / pw, oy / ph
                r2 = dx * dx + dy * dy
                if r2 == 0.0:
                    continue
                s = -self.G * (1.0 - math.exp(-r2 * self.k2)) / (total_mass * r2)
                K[0, ry, rx] = s * (0.0 if rx * 2 == mw else dx) / self.asp
                K[1, ry, rx] = s * (0.0 if ry * 2 == mh else dy)
        ikx = np.array([i if i <= mw // 2 else i - mw for i in range(mw)])
        iky = np.array([i if i <= mh // 2 else i - mh for i in range(mh)])
        aw = (np.sinc(iky / mh) ** 2)[:, None] * (np.sinc(ikx / mw) ** 2)[None, :]
        deconv = 1.0 / np.maximum(aw * aw, self.window_floor)
        return np.fft.fft2(K, axes=(1, 2)) * deconv

    def _image(self, d):
        return d - np.floor(d + 0.5)

    def _cic(self, p):
        gx, gy = p[0] * self.pw - 0.5, p[1] * self.ph - 0.5
        bx, by = int(math.floor(gx)), int(math.floor(gy))
        fx, fy = gx - bx, gy - by
        nodes = [(bx % self.mw, by % self.mh, (1 - fx) * (1 - fy)),
                 ((bx + 1) % self.mw, by % self.mh, fx * (1 - fy)),
                 (bx % self.mw, (by + 1) % self.mh, (1 - fx) * fy),
                 ((bx + 1) % self.mw, (by + 1) % self.mh, fx * fy)]
        return nodes, fx, fy

    def _home(self, p):
        return (int(math.floor(p[0] * self.pw)) % self.pw, int(math.floor(p[1] * self.ph)) % self.ph)

    def topology(self, x):
        cells = {}
        for i in range(len(x)):
            cells.setdefault(self._home(x[i]), []).append(i)
        for c in cells:
            cells[c].sort(reverse=True)
        return cells

    def _local(self, s, t, mass, total):
        px = self._image(s[0] - t[0]) * self.asp
        py = self._image(s[1] - t[1])
        r2 = px * px + py * py
        sc = self.G * mass * math.exp(-r2 * self.k2) / (total * max(r2, self.md2))
        return np.array([sc * px / self.asp, sc * py])

    def _stencil(self, hx, hy):
        for dy in range(-self.ry, self.ry + 1):
            for dx in range(-self.rx, self.rx + 1):
                yield dx, dy, ((hx + dx) % self.pw, (hy + dy) % self.ph)

    def accel(self, x, m, cells=None):
        x = np.asarray(x, np.float64)
        n = len(x)
        total = max(1.0, float(n))
        cells = self.topology(x) if cells is None else cells
        rho = np.zeros((self.mh, self.mw))
        for i in range(n):
            for a, b, w in self._cic(x[i])[0]:
                rho[b, a] += w * m[i]
        F = np.real(np.fft.ifft2(np.fft.fft2(rho)[None] * self._kernel(total), axes=(1, 2)))
        A = np.zeros((n, 2))
        for i in range(n):
            for a, b, w in self._cic(x[i])[0]:
                A[i] += w * F[:, b, a]
            for dx, dy, c in self._stencil(*self._home(x[i])):
                lst = cells.get(c, [])
                home = dx == 0 and dy == 0
                if not home and len(lst) > self.exact_limit:
                    centre = ((c[0] + 0.5) / self.pw, (c[1] + 0.5) / self.ph)
                    A[i] += self._local(centre, x[i], sum(m[j] for j in lst), total)
                else:
                    for j in lst[:self.home_limit if home else self.exact_limit]:
                        if j != i:
                            A[i] += self._local(x[j], x[i], m[j], total)
        return A, F

    def _term_vjp(self, px, py, mass, w, total):
        r2 = px * px + py * py
        e = math.exp(-r2 * self.k2)
        clamped = not (r2 > self.md2)
        q = 1.0 / (self.md2 if clamped else r2)
        phi = e * q
        dphi = -self.k2 * phi - (0.0 if clamped else phi * q)
        sw = w[0] * px / self.asp + w[1] * py
        cm = self.G / total * mass
        return (cm * (2.0 * dphi * px * sw + phi * w[0] / self.asp),
                cm * (2.0 * dphi * py * sw + phi * w[1]))

    def vjp(self, x, m, w, cells=None):
        """J_A(x)^T w, topology frozen at x."""
        x = np.asarray(x, np.float64)
        w = np.asarray(w, np.float64)
        n = len(x)
        total = max(1.0, float(n))
        cells = self.topology(x) if cells is None else cells
        Kh = self._kernel(total)
        _, F = self.accel(x, m, cells)
        g = np.zeros((n, 2))
        spread = np.zeros((self.mh, self.mw), complex)
        for i in range(n):
            nodes, fx, fy = self._cic(x[i])
            d = [w[i] @ F[:, b, a] for a, b, _ in nodes]
            g[i, 0] += self.pw * ((d[1] - d[0]) * (1 - fy) + (d[3] - d[2]) * fy)
            g[i, 1] += self.ph * ((d[2] - d[0]) * (1 - fx) + (d[3] - d[1]) * fx)
            for a, b, wt in nodes:
                spread[b, a] += wt * (w[i, 0] - 1j * w[i, 1])
            hx, hy = self._home(x[i])
            own = cells[(hx, hy)]
            count, rank = len(own), own.index(i)
            for dx, dy, c in self._stencil(hx, hy):
                lst = cells.get(c, [])
                if not lst:
                    continue
                home = dx == 0 and dy == 0
                if not home and len(lst) > self.exact_limit:
                    cpx = self._image((c[0] + 0.5) / self.pw - x[i, 0]) * self.asp
                    cpy = self._image((c[1] + 0.5) / self.ph - x[i, 1])
                    a, b = self._term_vjp(cpx, cpy, sum(m[j] for j in lst), w[i], total)
                    g[i] -= (self.asp * a, b)
                else:
                    for j in lst[:self.home_limit if home else self.exact_limit]:
                        if j == i:
                            continue
                        a, b = self._term_vjp(self._image(x[j, 0] - x[i, 0]) * self.asp,
                                              self._image(x[j, 1] - x[i, 1]), m[j], w[i], total)
                        g[i] -= (self.asp * a, b)
                seen = rank < self.home_limit if home else count <= self.exact_limit
                if not seen:
                    continue
                for k in lst:
                    if k == i:
                        continue
                    a, b = self._term_vjp(self._image(x[i, 0] - x[k, 0]) * self.asp,
                                          self._image(x[i, 1] - x[k, 1]), m[i], w[k], total)
                    g[i] += (self.asp * a, b)
        R = np.fft.ifft2(np.fft.fft2(spread) * (Kh[0] + 1j * Kh[1]))
        rbar = -R.real
        for i in range(n):
            nodes, fx, fy = self._cic(x[i])
            r = [rbar[b, a] for a, b, _ in nodes]
            g[i, 0] += m[i] * self.pw * ((r[1] - r[0]) * (1 - fy) + (r[3] - r[2]) * fy)
            g[i, 1] += m[i] * self.ph * ((r[2] - r[0]) * (1 - fx) + (r[3] - r[1]) * fx)
        return g
