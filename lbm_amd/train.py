"""Online training of the surrogate on trajectories of the batched LBM.

    python -m lbm_amd.train --iters 2000 --factor 2 --out surrogate.npy

A batch of square domains (obstacles, tau and inflow drawn like the batched
configuration: 1-12 circles/rectangles, tau in [0.52, 0.8], u_in in
[0.02, 0.1]) is spun up with the plain LBM, then snapshot pairs (t, t + 64) are
read from it continually (the renderer moments: three planes per domain);
each training step draws a minibatch from a ring of recent pairs.  The network
learns the 64-step change of (rho, u), pooled over F x F cells, in units of
S_RHO, S_U.  No reference counterpart.
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from .batch import Batch, gen_batch, gen_shapes
from .surrogate import S_RHO, S_U, UNet, fields_to_input, loss_weights, target_from


def velocity(jx, jy, rho):
    """(ux, uy) from momenta; cells with rho <= 0 (never in fluid) read 0."""
    safe = np.where(rho > 0, rho, 1.0)
    return np.where(rho > 0, jx / safe, 0.0), np.where(rho > 0, jy / safe, 0.0)


class PairSource:
    """Snapshot pairs (t, t + jump) from a batch of square domains."""

    def __init__(self, n_domains=16, size=256, factor=1, jump=64, spinup=1000, seed=7, device=0, tau=None,
                 u_in=None):
        params = gen_batch(seed, n_domains)
        self.factor = factor
        self.doms = [(size, size, 1.0 / tau if tau else om, u_in if u_in else u) for (_, _, om, u) in params]
        self.masks = [gen_shapes(size, size, seed * 1000 + i, 1, 12) for i in range(n_domains)]
        self.batch = Batch(self.doms, device=device)
        for i, m in enumerate(self.masks):
            self.batch.set_mask(i, m)
        self.jump = jump
        self.batch.step(spinup)
        self.prev = [self._fields(i) for i in range(n_domains)]

    def _fields(self, i):
        jx, jy, rho = self.batch.moments(i)
        return (rho, *velocity(jx, jy, rho))

    def advance(self):
        """Step every domain `jump` steps; returns one (x, y, w) per domain."""
        self.batch.step(self.jump)
        out = []
        for i, m in enumerate(self.masks):
            cur = self._fields(i)
            p = self.prev[i]
            om, u_in = self.doms[i][2], self.doms[i][3]
            x = fields_to_input(m, *p, 1.0 / om, u_in, self.factor)
            y = target_from(m, p, cur, self.factor)
            w = loss_weights(m, self.factor)
            out.append((x, y, w))
            self.prev[i] = cur
        return out

    def close(self):
        self.batch.close()


def train(iters=1000, minibatch=8, lr=1e-3, n_domains=16, size=256, factor=1, ring=256, seed=7, device=0,
          log_every=100, net=None, verbose=True, tau=None, u_in=None, advance_every=4):
    src = PairSource(n_domains=n_domains, size=size, factor=factor, seed=seed, device=device, tau=tau, u_in=u_in)
    tile = size // factor
    if net is None:
        net = UNet(minibatch, tile, tile, seed=seed, device=device)
    rng = np.random.default_rng(seed)
    pool = []
    losses = []
    t0 = time.perf_counter()
    for it in range(iters):
        if it % advance_every == 0 or len(pool) < minibatch:
            pool.extend(src.advance())
            pool = pool[-ring:]
        pick = rng.choice(len(pool), minibatch, replace=len(pool) < minibatch)
        x = np.stack([pool[j][0] for j in pick])
        y = np.stack([pool[j][1] for j in pick])
        w = np.stack([pool[j][2] for j in pick])
        cur_lr = lr * 0.5 * (1 + np.cos(np.pi * it / max(1, iters)))  # cosine decay
        losses.append(net.train_step(x, y, w, float(cur_lr)))
        if verbose and (it % log_every == 0 or it == iters - 1):
            print(f"[train] iter {it} loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f} s)", flush=True)
    src.close()
    return net, np.array(losses)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=1000)
    ap.add_argument("--minibatch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--domains", type=int, default=16)
    ap.add_argument("--factor", type=int, default=1, choices=(1, 2))
    ap.add_argument("--size", type=int, default=0, help="domain side (default 256 x factor)")
    ap.add_argument("--out", default="surrogate.npy")
    a = ap.parse_args()
    size = a.size or 256 * a.factor
    net, losses = train(a.iters, a.minibatch, a.lr, a.domains, size, a.factor)
    np.save(a.out, net.get_params())
    print(f"saved {a.out}; loss {losses[:10].mean():.4e} -> {losses[-10:].mean():.4e}; "
          f"S_RHO={S_RHO} S_U={S_U} factor={a.factor}")


if __name__ == "__main__":
    main()
