"""Training data, GPU training and use of the learned BZ-mesh proposal.

The method generates its own training data: the probe proposal of the
windowed mesh (``bzmesh.WindowedMesh``, mode "probe") computes the exact
indicator e_l(c) on nested meshes, and every active cell of every level gives
one sample (level-l features -> log10 e_l(c)/|I|).  The systems are seeded
random layer systems of the configs' generation law (Hermitian site blocks,
G_r(R) ~ exp(-2|R|)), smaller than config 1 so that 200 of them fit in
minutes of GPU time, with energies near the real axis.  The network is the
3-layer MLP of libskkr_mlp.so, trained on the GPU by its own kernels.

CLI (GPU):
    python -m skkr_budapest_amd.bztrain generate --systems 200 --out data.npz
    python -m skkr_budapest_amd.bztrain train --data data.npz --out model.npz
"""
from __future__ import annotations

import argparse
import math
import os
import time

import numpy as np

from . import bzmesh as BM
from . import systems as S

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "data", "bzmesh_mlp.npz")
LABEL_CLIP = (-14.0, 0.0)
THETA = 1e-8          # indicator threshold, relative to |I|
MARGIN = 0.5          # decades subtracted from log10 THETA for the prediction
SEED0 = 7100


def random_system(i: int):
    """The i-th seeded training system and its energy."""
    rng = np.random.default_rng(SEED0 + i)
    lk = int(rng.choice([4, 8]))
    s = int(rng.choice([2, 4]))
    P = int(rng.choice([2, 3, 4]))
    r2max = int(rng.choice([4, 8]))
    im_e = float(10.0 ** rng.uniform(math.log10(0.003), math.log10(0.03)))
    E = complex(rng.uniform(-0.4, 0.8), im_e)
    cfg = S.small_config(lk=lk, s=s, P=P, r2max=r2max, tmodel="herm", cfg_id=SEED0 + i,
                         energies=(E,), mesh=4, name=f"bztrain-{i}")
    return S.make_system(cfg), E


def trace_evaluator(ctx, E, chunk: int = 65536):
    """f(k) = sum over sites of tr tau(k) from the GPU k-loop."""
    def ev(k):
        k = np.ascontiguousarray(k, np.float64)
        out = np.empty(len(k), np.complex128)
        for a in range(0, len(k), chunk):
            b = min(a + chunk, len(k))
            r = ctx.tau_bz(E, k[a:b], np.zeros(b - a))
            out[a:b] = r["trk"].sum(1)
        return out
    return ev


class Recorder:
    """Collects (features, label) of active cells, at most `cap` per level."""

    def __init__(self, cap: int, seed: int):
        self.cap, self.rng = cap, np.random.default_rng(seed)
        self.x, self.y, self.lv = [], [], []

    def __call__(self, level, feats, labels, active):
        idx = np.nonzero(active)[0]
        if len(idx) > self.cap:
            idx = self.rng.choice(idx, self.cap, replace=False)
        self.x.append(feats[idx])
        self.y.append(np.clip(labels[idx], *LABEL_CLIP).astype(np.float32))
        self.lv.append(np.full(len(idx), level, np.int16))


def generate(n_systems: int, n0: int = 4, lmax: int = 8, theta: float = 1e-8,
             cap: int = 1500, device: int = 0, max_seconds: float = None, log=print):
    from . import gpu
    rec = Recorder(cap, seed=SEED0)
    sys_id = []
    t0 = time.time()
    evals = 0
    for i in range(n_systems):
        if max_seconds is not None and time.time() - t0 > max_seconds:
            log(f"stopping after {i} systems (time limit)", flush=True)
            break
        sy, E = random_system(i)
        with gpu.Context.for_system(sy, device=device) as ctx:
            ctx.set_tinv(*sy.tinv(E))
            m = BM.WindowedMesh(trace_evaluator(ctx, E), n0=n0, lmax=lmax, im_e=E.imag)
            n_before = len(rec.y)
            r = m.integrate("probe", theta=theta, record=rec)
        evals += r["n_evals"]
        sys_id.extend([i] * (len(rec.y) - n_before))
        log(f"system {i}: lk={sy.cfg.lk} s={sy.cfg.s} P={sy.cfg.P} E={E:.4f} "
            f"levels {r['levels']} evals {r['n_evals']} flagged {r['flagged']} "
            f"({time.time() - t0:.1f} s)", flush=True)
    sid = np.concatenate([np.full(len(y), k, np.int32) for y, k in zip(rec.y, sys_id)])
    return {"x": np.concatenate(rec.x), "y": np.concatenate(rec.y),
            "level": np.concatenate(rec.lv), "system": sid, "evals": evals}


def train(x, y, steps: int = 3000, lr: float = 3e-3, seed: int = 1, device: int = 0):
    from .mlp import Mlp
    m = Mlp(x.shape[1], seed=seed, device=device)
    hist = m.fit(x, y, steps=steps, lr=lr)
    return m, hist


def save_model(path, m):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=m.params(), shift=m.shift, scale=m.scale,
             features=np.array(BM.FEATURES))


def load_model(path=MODEL_PATH, device: int = 0):
    """The trained proposal network on the GPU (libskkr_mlp.so)."""
    from .mlp import Mlp
    d = np.load(path, allow_pickle=False)
    if tuple(d["features"]) != BM.FEATURES:
        raise ValueError("model features differ from bzmesh.FEATURES")
    m = Mlp(len(BM.FEATURES), device=device)
    m.set_params(d["params"])
    m.shift, m.scale = d["shift"].astype(np.float32), d["scale"].astype(np.float32)
    return m


def integrate_bz(evaluate, n0: int, lmax: int, im_e: float, tol: float = 1e-6, model=None,
                 theta: float = THETA, margin: float = MARGIN):
    """BZ average of f with the learned proposal, or -- without a model (none
    trained, or trained on other features) -- the fallback: plain uniform
    refinement until two successive levels agree to tol."""
    if model is None:
        hist = BM.uniform_levels(evaluate, n0, lmax, tol=tol)
        converged = len(hist) > 1 and abs(hist[-1][0] - hist[-2][0]) <= tol * abs(hist[-1][0])
        return {"integral": hist[-1][0], "n_evals": hist[-1][1], "method": "uniform",
                "converged": converged}
    m = BM.WindowedMesh(evaluate, n0=n0, lmax=lmax, im_e=im_e)
    r = m.integrate("learned", theta=theta, predict=model.predict, margin=margin)
    r["method"] = "learned"
    return r


def compare(evaluate, model, n0: int, lmax: int, im_e: float, tol: float = 1e-6,
            theta: float = THETA, margin: float = MARGIN):
    """Learned proposal, probe proposal and plain uniform refinement on one
    integrand.  The judge is the uniform level-lmax integral: every k-point of
    the nested lattice is evaluated once, the other runs read those values
    and count the distinct points they use."""
    hist, grid = BM.uniform_levels(evaluate, n0, lmax, return_grid=True)
    ref = hist[-1][0]
    M = grid.shape[0]

    def lookup(k):
        ix = np.rint((k[:, 0] + math.pi) * M / (2 * math.pi)).astype(np.int64) % M
        iy = np.rint((k[:, 1] + math.pi) * M / (2 * math.pi)).astype(np.int64) % M
        return grid[ix, iy]

    rel = lambda x: abs(x - ref) / abs(ref)
    lr = BM.WindowedMesh(lookup, n0, lmax, im_e).integrate("learned", theta=theta,
                                                           predict=model.predict, margin=margin)
    pr = BM.WindowedMesh(lookup, n0, lmax, im_e).integrate("probe", theta=theta)
    need = next((n for I, n in hist[:-1] if rel(I) <= tol), None)
    return {"uniform": [(n, rel(I)) for I, n in hist], "uniform_needed": need,
            "learned": (lr["n_evals"], rel(lr["integral"])),
            "probe": (pr["n_evals"], rel(pr["integral"])), "reference": ref}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--systems", type=int, default=200)
    g.add_argument("--n0", type=int, default=4)
    g.add_argument("--lmax", type=int, default=8)
    g.add_argument("--theta", type=float, default=1e-8)
    g.add_argument("--cap", type=int, default=1500)
    g.add_argument("--max-seconds", type=float, default=None)
    g.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--first", type=int, default=1000)
    e.add_argument("--count", type=int, default=10)
    e.add_argument("--n0", type=int, default=8)
    e.add_argument("--lmax", type=int, default=7)
    e.add_argument("--model", default=MODEL_PATH)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--out", default=MODEL_PATH)
    a = ap.parse_args(argv)
    if a.cmd == "evaluate":
        from . import gpu
        model = load_model(a.model)
        for i in range(a.first, a.first + a.count):
            sy, E = random_system(i)
            with gpu.Context.for_system(sy) as ctx:
                ctx.set_tinv(*sy.tinv(E))
                r = compare(trace_evaluator(ctx, E), model, a.n0, a.lmax, E.imag)
            print(f"system {i} lk={sy.cfg.lk} s={sy.cfg.s} P={sy.cfg.P} E={E:.4f}: "
                  f"uniform {[(n, float(f'{x:.2e}')) for n, x in r['uniform']]} "
                  f"needed {r['uniform_needed']} | learned {r['learned'][0]} "
                  f"err {r['learned'][1]:.2e} | probe {r['probe'][0]} err {r['probe'][1]:.2e}",
                  flush=True)
        return
    if a.cmd == "generate":
        d = generate(a.systems, a.n0, a.lmax, a.theta, a.cap, max_seconds=a.max_seconds)
        np.savez_compressed(a.out, **d)
        print(f"wrote {a.out}: {len(d['y'])} samples, {d['evals']} k-evaluations")
    else:
        d = np.load(a.data, allow_pickle=False)
        m, hist = train(d["x"], d["y"], a.steps, a.lr)
        save_model(a.out, m)
        print(f"trained: loss {hist[0]:.4f} -> {hist[-1]:.4f}; wrote {a.out}")


if __name__ == "__main__":
    main()
