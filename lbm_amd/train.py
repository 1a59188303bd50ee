"""Online training of the surrogate on trajectories of the batched LBM.

    python -m lbm_amd.train --iters 2000 --out surrogate.npy

A batch of 256 x 256 domains (obstacles, tau and inflow drawn like the batched
configuration: 1-12 circles/rectangles, tau in [0.52, 0.8], u_in in
[0.02, 0.1]) is spun up with the plain LBM, then snapshot pairs (t, t + 64) are
read from it continually; each training step draws a minibatch from a ring of
recent pairs.  The network learns the 64-step change of (rho, u) in units of
S_RHO, S_U.  No reference counterpart.
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from .batch import Batch, gen_batch, gen_shapes
from .surrogate import S_RHO, S_U, UNet, fields_to_input, target_from

CX = np.array([-1, 0, 1, -1, 0, 1, -1, 0, 1], np.float32)
CY = np.array([-1, -1, -1, 0, 0, 0, 1, 1, 1], np.float32)


def macro(f9):
    """(rho, ux, uy) of a (9, H, W) state; uy is north-positive."""
    rho = f9.sum(0)
    ux = np.tensordot(CX, f9, 1) / rho
    uy = -np.tensordot(CY, f9, 1) / rho
    return rho, ux, uy


class PairSource:
    """Snapshot pairs (t, t + jump) from a batch of square domains."""

    def __init__(self, n_domains=16, size=256, jump=64, spinup=1000, seed=7, device=0):
        params = gen_batch(seed, n_domains)
        self.doms = [(size, size, om, u) for (_, _, om, u) in params]
        self.masks = [gen_shapes(size, size, seed * 1000 + i, 1, 12) for i in range(n_domains)]
        self.batch = Batch(self.doms, device=device)
        for i, m in enumerate(self.masks):
            self.batch.set_mask(i, m)
        self.jump = jump
        self.batch.step(spinup)
        self.prev = [macro(self.batch.read_f(i)) for i in range(n_domains)]

    def advance(self):
        """Step every domain `jump` steps; returns one (x, y, w) per domain."""
        self.batch.step(self.jump)
        out = []
        for i, m in enumerate(self.masks):
            cur = macro(self.batch.read_f(i))
            p = self.prev[i]
            x = fields_to_input(m, *p)
            y = target_from(m, p, cur)
            w = (m != 1).astype(np.float32)
            out.append((x, y, w))
            self.prev[i] = cur
        return out


def train(iters=1000, minibatch=8, lr=1e-3, n_domains=16, ring=256, seed=7, device=0, log_every=100,
          net=None, verbose=True):
    src = PairSource(n_domains=n_domains, seed=seed, device=device)
    size = src.doms[0][0]
    if net is None:
        net = UNet(minibatch, size, size, seed=seed, device=device)
    rng = np.random.default_rng(seed)
    pool = []
    losses = []
    t0 = time.perf_counter()
    for it in range(iters):
        if it % 4 == 0 or len(pool) < minibatch:
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
    src.batch.close()
    return net, np.array(losses)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=1000)
    ap.add_argument("--minibatch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--domains", type=int, default=16)
    ap.add_argument("--out", default="surrogate.npy")
    a = ap.parse_args()
    net, losses = train(a.iters, a.minibatch, a.lr, a.domains)
    np.save(a.out, net.get_params())
    print(f"saved {a.out}; loss {losses[:10].mean():.4e} -> {losses[-10:].mean():.4e}; "
          f"S_RHO={S_RHO} S_U={S_U}")


if __name__ == "__main__":
    main()
