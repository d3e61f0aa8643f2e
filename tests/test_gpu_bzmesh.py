"""The learned BZ-mesh proposal on the GPU: the trained network
(skkr_budapest_amd/data/bzmesh_mlp.npz, HIP inference) against the exact
indicator on a held-out system, and its gain over the method without it
(the probe proposal) and over plain uniform refinement, which stays the judge.
Integrals are checked against the closed form of the square-lattice Green's
function (test_bzmesh.py) or against uniform refinement of the same k-lattice."""
import math

import numpy as np
import pytest
from scipy.special import ellipk
from scipy.stats import spearmanr

pytestmark = pytest.mark.gpu


def _model():
    from skkr_budapest_amd import bztrain as BT
    return BT.load_model()


def test_trained_network_predicts_the_indicator_on_a_held_out_system():
    from skkr_budapest_amd import bzmesh as BM
    from skkr_budapest_amd import bztrain as BT
    from skkr_budapest_amd import gpu
    sy, E = BT.random_system(1000)           # systems 0..199 are the training set
    print(f"bzmesh: held-out system 1000, E={E}", flush=True)
    rec = BT.Recorder(cap=100000, seed=0)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        m = BM.WindowedMesh(BT.trace_evaluator(ctx, E), n0=8, lmax=6, im_e=E.imag)
        m.integrate("probe", theta=BT.THETA, record=rec)
    x, y = np.concatenate(rec.x), np.concatenate(rec.y)
    pred = _model().predict(x)
    rms = float(np.sqrt(np.mean((pred - y) ** 2)))
    rho = spearmanr(pred, y).statistic
    print(f"bzmesh: {len(y)} cells, rms {rms:.2f} decades, rank correlation {rho:.3f}", flush=True)
    assert rms < 1.0 and rho > 0.85


def test_learned_mesh_gain_on_a_localized_peak():
    """z = 2.002: the lattice Green's function peaks at k = 0 (width ~0.06).
    Plain uniform refinement from 16 x 16 needs 256 x 256 points for 1e-6;
    the learned proposal reaches it with a third of that or less, and with
    fewer points than the probe proposal it replaces."""
    from skkr_budapest_amd import bzmesh as BM
    from skkr_budapest_amd import bztrain as BT
    z = 2.002
    exact = 2.0 / (math.pi * z) * ellipk(4.0 / z ** 2)
    ev = lambda k: 1.0 / (z - np.cos(k[:, 0]) - np.cos(k[:, 1])) + 0j
    hist = BM.uniform_levels(ev, 16, 6)
    need = next(n for I, n in hist if abs(I - exact) <= 1e-6 * exact)
    r = BT.integrate_bz(ev, 16, 6, im_e=0.005, model=_model())
    p = BM.WindowedMesh(ev, 16, 6, 0.005).integrate("probe", theta=BT.THETA)
    print(f"bzmesh: uniform {need}, learned {r['n_evals']} "
          f"(err {abs(r['integral'] - exact) / exact:.1e}), probe {p['n_evals']}", flush=True)
    assert abs(r["integral"] - exact) <= 1e-6 * exact
    assert 3 * r["n_evals"] <= need
    assert r["n_evals"] < p["n_evals"]


def test_learned_mesh_on_held_out_layer_systems():
    """tr G of three held-out layer systems through the GPU k-loop, judged by
    uniform refinement to level 7 of the 8 x 8 start (1024 x 1024).  The
    indicator threshold (1e-8 of |I| per cell, 0.5 decades of margin on the
    prediction) puts the learned result at ~1e-6 (4e-7 ... 1.8e-6 measured
    on eight held-out systems); the learned proposal uses fewer than half the
    k-evaluations of the probe proposal it replaces."""
    from skkr_budapest_amd import bztrain as BT
    from skkr_budapest_amd import gpu
    model = _model()
    learned = probe = 0
    for i in (1000, 1001, 1002):
        sy, E = BT.random_system(i)
        print(f"bzmesh: held-out system {i}, E={E}", flush=True)
        with gpu.Context.for_system(sy) as ctx:
            ctx.set_tinv(*sy.tinv(E))
            r = BT.compare(BT.trace_evaluator(ctx, E), model, 8, 7, E.imag)
        print(f"bzmesh: uniform {r['uniform']}, learned {r['learned']}, probe {r['probe']}",
              flush=True)
        assert r["learned"][1] <= 2e-6
        learned += r["learned"][0]
        probe += r["probe"][0]
    assert 2 * learned < probe


def test_learned_mesh_on_the_config4_system():
    """The config-4 system itself (lk = 32, 4 sites per PL, n = 128, 20 PLs,
    E = 0.6 + 0.005i) through the GPU k-loop, at 2e-5 relative error from the
    32 x 32 triangulation with up to 4 levels (512 x 512).  The judge is
    uniform level 5 (1024 x 1024), a level finer than anything judged and
    evaluated on its own; its error (~1e-7 against level 6 in the full
    benchmark) is far below the tolerance.  Uniform refinement first reaches
    2e-5 at level 4 (262,144 points: level 3 is at 2.5e-5); the learned
    proposal, with the indicator threshold raised tenfold for the looser
    tolerance, needs about half (130,944 points measured, error 1.8e-5).
    At coarser settings (1e-4, up to level 3) the proposal refines nearly
    every cell and gains nothing: the gain of the learned mesh is in the deep
    levels."""
    from skkr_budapest_amd import bzmesh as BM
    from skkr_budapest_amd import bztrain as BT
    from skkr_budapest_amd import gpu
    from skkr_budapest_amd import systems as S
    cfg = S.CONFIGS[4]
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    tol, n0, lmax = 2e-5, 32, 4
    print(f"bzmesh: config 4 system n={cfg.n} P={cfg.P} E={E}", flush=True)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        ev = BT.trace_evaluator(ctx, E)
        lr = BM.WindowedMesh(ev, n0, lmax, E.imag).integrate(
            "learned", theta=10 * BT.THETA, predict=_model().predict, margin=BT.MARGIN)
        print(f"bzmesh: learned {lr['n_evals']} points", flush=True)
        judge = BM.uniform_levels(ev, n0, lmax + 1)
        print(f"bzmesh: judge {judge[-1][1]} points", flush=True)
    ref = judge[-1][0]
    rel = lambda x: abs(x - ref) / abs(ref)
    errs = [rel(I) for I, _ in judge[:-1]]
    print(f"bzmesh: uniform errors {['%.2e' % e for e in errs]}, learned err {rel(lr['integral']):.2e}",
          flush=True)
    assert errs[3] > tol >= errs[4]        # uniform needs level 4
    need = judge[4][1]
    assert rel(lr["integral"]) <= tol
    assert lr["n_evals"] <= 0.6 * need
