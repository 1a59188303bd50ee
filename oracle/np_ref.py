"""Second f32 restatement of the reference step, in numpy — TEST INFRASTRUCTURE ONLY.

Written independently of refcpu.c to cross-check it bit for bit on small grids.
Where refcpu.c loops pass by pass over plane pairs, this version gathers all
eight moving populations with one table of flattened-index offsets and applies
the whole collision with vector expressions.  numpy float32 arithmetic rounds
each operation to f32 and never fuses a multiply-add, so the expression trees
below reproduce the WGSL ones exactly:

  moments ............ pre_collision/corner_pre_collision.wgsl:19-21,
                       pre_collision/cardinal_pre_collision.wgsl:19-21
  relaxation ......... collision/corner_collision.wgsl:24-42,
                       collision/cardinal_collision.wgsl:25-43
  streaming .......... stream/{e_w,n_s,ne_sw,se_nw}_stream.wgsl:31-62
"""
from __future__ import annotations

import numpy as np

F = np.float32
# plane k -> flattened upstream offset as a multiple of (1, X); lbm-wgpu src/lbm.rs:632-640
# (upstream of population k at index i is i - c_k).
_CX = (-1, 0, 1, -1, 0, 1, -1, 0, 1)
_CY = (-1, -1, -1, 0, 0, 0, 1, 1, 1)
_OPP = (8, 7, 6, 5, 4, 3, 2, 1, 0)


class NpRef:
    def __init__(self, nx: int, ny: int, omega: float):
        self.X, self.Y = nx, ny
        self.T = nx * ny
        self.omega = F(omega)
        self.sets = np.zeros((2, 9, self.T), F)
        self.mask = np.zeros(self.T, np.uint32)
        self.mask[:nx] = 1
        self.mask[(ny - 1) * nx:] = 1
        self.dens = np.zeros((3, self.T), F)
        self.step_n = 0

    def fill(self, f9):
        f9 = np.asarray(f9, F).reshape(9, -1)
        if f9.shape[1] == 1:
            f9 = np.repeat(f9, self.T, axis=1)
        self.sets[0] = f9
        self.sets[1] = f9

    def _moments(self, p):
        s = self.sets[p]
        nw, n, ne, w, e, sw, st, se = s[0], s[1], s[2], s[3], s[5], s[6], s[7], s[8]
        jx = (((ne + se) - nw) - sw) + (e - w)
        jy = (((ne + nw) - se) - sw) + (n - st)
        rho = (((ne + se) + nw) + sw) + (((e + n) + st) + w)
        return jx, jy, rho

    def collide(self):
        p = self.step_n % 2
        jx, jy, rho = self._moments(p)
        s = self.sets[p]
        rest = self.sets[0][4]
        rho = rho + rest
        self.dens[:] = (jx, jy, rho)
        ux = jx / rho
        uy = jy / rho
        ux3, uy3 = F(3) * ux, F(3) * uy
        ux2, uy2 = ux * ux, uy * uy
        uxuy2 = (F(2) * ux) * uy
        u2 = ux2 + uy2
        u215 = F(1.5) * u2
        r36 = F(1.0 / 36.0) * rho
        r9 = F(1.0 / 9.0) * rho
        r49 = F(4.0 / 9.0) * rho
        om = self.omega
        eq = {
            2: r36 * ((((F(1) + ux3) + uy3) + F(4.5) * (u2 + uxuy2)) - u215),
            8: r36 * ((((F(1) + ux3) - uy3) + F(4.5) * (u2 - uxuy2)) - u215),
            0: r36 * ((((F(1) - ux3) + uy3) + F(4.5) * (u2 - uxuy2)) - u215),
            6: r36 * ((((F(1) - ux3) - uy3) + F(4.5) * (u2 + uxuy2)) - u215),
            5: r9 * (((F(1) + ux3) + F(4.5) * ux2) - u215),
            3: r9 * (((F(1) - ux3) + F(4.5) * ux2) - u215),
            1: r9 * (((F(1) + uy3) + F(4.5) * uy2) - u215),
            7: r9 * (((F(1) - uy3) + F(4.5) * uy2) - u215),
        }
        for k, fe in eq.items():
            s[k] = s[k] + om * (fe - s[k])
        self.sets[0][4] = rest + om * (r49 * (F(1) - u215) - rest)

    def stream(self):
        p = self.step_n % 2
        X, T = self.X, self.T
        src, dst = self.sets[p], self.sets[1 - p]
        idx = np.arange(T, dtype=np.int64)
        live = (self.mask != 1) & (idx % X != 0) & (idx // X < self.Y - 1)
        cells = idx[live]
        for k in range(9):
            if k == 4:
                continue
            up = cells - (_CX[k] + _CY[k] * X)   # flattened, no x wrap check
            inb = (up >= 0) & (up < T)
            solid = np.ones(cells.size, bool)
            solid[inb] = self.mask[up[inb]] == 1
            vals = np.where(solid, src[_OPP[k]][cells], src[k][np.clip(up, 0, T - 1)])
            dst[k][cells] = vals

    def step(self, n=1):
        for _ in range(n):
            self.collide()
            self.stream()
            self.step_n += 1

    def state(self):
        p = self.step_n % 2
        out = self.sets[p].copy()
        out[4] = self.sets[0][4]
        return out.reshape(9, self.Y, self.X)
