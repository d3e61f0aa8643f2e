"""Benchmark of the per-(E,k) layer kernel on one MI355X.

A step is one pass of the hot path over one energy of the workload: the
k-loop body for every k-point of the mesh (lattice Fourier sum, decimation of
both leads, Godfrin sweeps, BZ accumulation), inputs resident in HBM.

Default workload: the relativistic layer system with lk=32, 4 sites per
principal layer (n=128), 20 interface PLs, 25 in-plane vectors, one energy
E=0.6+0.02i Ry and a 64x64 k-mesh (4096 k-points), timed first in FP64 and
then in the two mixed modes (INT8 digit products, split-BF16 products);
`value` is the INT8 mixed mode's k-points/s, `config` holds every rate, the
ratios to FP64 and each mixed result's error against the FP64 one (tau_BZ,
tr G, DOS).  --fp64-only times FP64 alone.

--kkr runs the k-loop on the reference's own structure constants instead of
the synthetic Fourier sum: the Nb(110) stack of sample/02-layer-scf (lmax 2,
4 layers per principal layer, 3 interface PLs), Kambe/direct sums and the
screening per k on the GPU, E = 0.5+0.05i, 64x64 mesh of the reciprocal cell.

--config 2 / 3 run a 16-point contour instead (a step is the whole contour):
config 2 the same system with a 64x64 mesh per energy, config 3 the BdG
system (lk=64, n=256, 12 PLs) with 256 ... 16384 k per energy, geometric in
Im E.  --config 0 is the small nonrelativistic system (16x16 mesh).

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("BZ-integrated tau k-points/s per energy: mixed+refine vs FP64, "
          "at 1e-8 rel err")
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0  # dense BF16 MFMA, MI355X datasheet
PEAK_INT8_TOPS = 5000.0    # dense INT8 MFMA, MI355X datasheet
# HBM bytes moved per algorithmic byte by zgemm_f64_kernel<128,128,16>: rocprofv3
# FETCH_SIZE + WRITE_SIZE of 1024 x complex 128^3 products, calibrated with a
# plain 256 MiB copy (profiles/hbm_traffic_pmc.md): 805.8 MB / 805.3 MB
GEMM_TRAFFIC_PER_ALGO_BYTE = 805.8 / 805.3


def algorithmic_flops_per_k(n, P, it_l, it_r):
    """(56P-48) n^3 for the sweeps + (56 I + 24) n^3 per lead."""
    return ((56 * P - 48) + (56 * it_l + 24) + (56 * it_r + 24)) * float(n) ** 3


# k-points/s of the reference's own compiled k-loop (flang -O2, 8 processes,
# round-robin k) over those of the C oracle (8 threads) on the same cores, for
# the n = 64 stack, the largest the reference's param.h admits
# (profiles/cpu_reference_vs_port_cfg0_n64.md, oracle/ref/time_cfg0.py)
REFERENCE_OVER_PORT = 1.20


def cpu_baseline(sy, E, k, w):
    import oracle
    nworkers = min(16, os.cpu_count() or 1)
    nk = 4 * nworkers
    idx = np.linspace(0, len(w) - 1, nk).astype(int)
    t0 = time.perf_counter()
    oracle.tau_bz(sy, E, k[idx], w[idx], nworkers=nworkers)
    dt = time.perf_counter() - t0
    return {"value": nk / dt, "unit": "k-points/s", "cores": nworkers, "kind": "port",
            "sample": f"{nk} k-points of the same energy and mesh, FP64 C oracle, "
                      f"{nworkers} threads, round-robin k split, {dt:.1f} s",
            "reference_over_port": REFERENCE_OVER_PORT,
            "reference_equivalent_value": REFERENCE_OVER_PORT * nk / dt}


def contour_bench(args, cfg, sy):
    """A step is the whole contour.  Config 2 times FP64 and, unless
    --fp64-only, the two mixed modes; `value` is the INT8 mixed mode's rate and
    every energy reports the mixed results' errors against FP64 (tau_BZ,
    tr G, DOS).  Config 3 (BdG, n = 256) times FP64, or the modes --modes
    names."""
    from skkr_budapest_amd import contour, energy_loop, gpu
    from skkr_budapest_amd import observables as O
    z, wz = contour.config_contour()
    if args.config == 2:
        sides = [cfg.mesh] * len(z)
    else:
        sides = energy_loop.geometric_mesh_sides(z.imag, 256, 16384)
    nk_total = sum(s * s for s in sides)
    print(f"bench: config {args.config} contour {len(z)} energies, {nk_total} k in total",
          file=sys.stderr, flush=True)
    modes = [("fp64", gpu.FP64)]
    if (args.config == 2 or args.modes) and not args.fp64_only:
        modes += [("mixed_int8", gpu.MIXED_INT8), ("mixed_bf16", gpu.MIXED_BF16)]
    if args.modes:
        want = args.modes.split(",")
        modes = [md for md in modes if md[0] in want]
    runs = {}
    for name, prec in modes:
        ctx = gpu.Context.for_system(sy, precision=prec)
        for i in range(args.warmup):
            energy_loop.run_contour(ctx, sy, z[-1:], wz[-1:], sides[-1:])
        times = []
        recs = total = None
        for i in range(args.steps):
            t0 = time.perf_counter()
            recs, total = energy_loop.run_contour(ctx, sy, z, wz, sides)
            times.append(time.perf_counter() - t0)
            print(f"bench: {name} contour step {i} {times[-1]:.2f} s, integral of tr G {total}",
                  file=sys.stderr, flush=True)
        ctx.close()
        runs[name] = {"sec": float(np.mean(times)), "recs": recs, "total": total}
    head_name = "mixed_int8" if "mixed_int8" in runs else "fp64"
    head = runs[head_name]
    fp = runs.get("fp64", head)
    energies = []
    for j, r in enumerate(head["recs"]):
        e = {"E": [r["E"].real, r["E"].imag], "nk": r["nk"], "seconds": r["seconds"],
             "dos": r["dos"], "mean_iter": float(r["iters"].mean())}
        for name in runs:
            if name != "fp64" and "fp64" in runs:
                e[f"{name}_error_vs_fp64"] = O.error_report(runs[name]["recs"][j]["tau_bz"],
                                                            fp["recs"][j]["tau_bz"])
        energies.append(e)
    cfg_out = {"workload": f"cfg{args.config}: lk={cfg.lk} s={cfg.s} P={cfg.P} n={cfg.n} "
                           f"N_R={cfg.n_R}, 16-point contour E_b=-0.8 E_F=0.6, "
                           f"{nk_total} k in total",
               "precision": head_name,
               "integrated_trG": [head["total"].real, head["total"].imag],
               "energies": energies}
    for name, r in runs.items():
        cfg_out[name] = {"k_points_per_s": nk_total / r["sec"], "seconds_per_contour": r["sec"]}
        if name != "fp64" and "fp64" in runs:
            cfg_out[name]["over_fp64"] = fp["sec"] / r["sec"]
            cfg_out[name]["integrated_trG_rel_err"] = abs(r["total"] - fp["total"]) / abs(fp["total"])
            cfg_out[name]["max_tau_rel_err"] = max(e[f"{name}_error_vs_fp64"]["tau_rel_err"]
                                                   for e in energies)
    out = {
        "metric": METRIC, "value": nk_total / head["sec"], "unit": "k-points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * head["sec"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (seeded G_r(R), t^-1)",
        "config": cfg_out,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def mesh_bench(args, cfg, sy):
    """Config 4: k-evaluations of the learned adaptive BZ mesh against plain
    uniform refinement, both from the 32 x 32 triangulation and both allowed
    up to level 5 (1024 x 1024), for 1e-6 relative error in the BZ integral
    of tr G at E = 0.6 + 0.005i.  The judge is uniform level 6 (2048 x 2048),
    one level finer than anything judged; its own error is bounded by the
    level-5/level-6 difference.  The method without the network (probe
    proposal) is counted on the same k-points."""
    from skkr_budapest_amd import bzmesh as BM
    from skkr_budapest_amd import bztrain as BT
    from skkr_budapest_amd import gpu
    E = cfg.energies[0]
    n0, lmax, tol = 32, 5, 1e-6
    ctx = gpu.Context.for_system(sy)
    ctx.set_tinv(*sy.tinv(E))
    ev0 = BT.trace_evaluator(ctx, E)
    model = BT.load_model()
    ev0(np.zeros((64, 2)))                                  # warm-up
    t0 = time.perf_counter()
    lm = BM.WindowedMesh(ev0, n0, lmax, E.imag)
    lr = lm.integrate("learned", theta=BT.THETA, predict=model.predict, margin=BT.MARGIN)
    t_learned = time.perf_counter() - t0
    print(f"bench: learned mesh {lr['n_evals']} k in {t_learned:.1f} s, flagged {lr['flagged']}",
          file=sys.stderr, flush=True)
    chunk = 1 << 18

    def ev(k):
        out = np.empty(len(k), np.complex128)
        for a in range(0, len(k), chunk):
            out[a:a + chunk] = ev0(k[a:a + chunk])
            print(f"bench: uniform points {min(a + chunk, len(k))} of {len(k)} at this level",
                  file=sys.stderr, flush=True)
        return out

    t0 = time.perf_counter()
    hist, grid = BM.uniform_levels(ev, n0, lmax + 1, return_grid=True)
    t_uniform = time.perf_counter() - t0
    ref = hist[-1][0]
    ctx.close()
    finest = grid.shape[0]

    def lookup(k):
        """f at the uniform level-(lmax+1) points, which hold the whole nested lattice."""
        ix = np.rint((k[:, 0] + np.pi) * finest / (2 * np.pi)).astype(int) % finest
        iy = np.rint((k[:, 1] + np.pi) * finest / (2 * np.pi)).astype(int) % finest
        return grid[ix, iy]

    pm = BM.WindowedMesh(lookup, n0, lmax, E.imag)
    pr = pm.integrate("probe", theta=BT.THETA)
    rel = lambda x: abs(x - ref) / abs(ref)
    levels = [{"level": l, "k_evals": n, "rel_err_vs_judge": rel(I)} for l, (I, n) in
              enumerate(hist[:-1])]
    # uniform refinement is judged on the levels it may use (0..lmax), never on the judge
    uni_need = next((n for I, n in hist[:-1] if rel(I) <= tol), None)
    out = {
        "metric": "k-evaluations to 1e-6 rel err in the BZ integral of tr G: "
                  "uniform refinement / learned mesh",
        "value": (uni_need / lr["n_evals"]) if uni_need else None, "unit": "x fewer k-evaluations",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * t_learned,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (seeded G_r(R), t^-1); MLP trained on 200 seeded systems",
        "config": {"workload": f"cfg4: lk={cfg.lk} s={cfg.s} P={cfg.P} n={cfg.n} E={E}, "
                               f"triangulation from {n0}x{n0}, up to {lmax} levels",
                   "tol": tol, "judge": f"uniform level {lmax + 1} ({finest}x{finest}, "
                                        f"{hist[-1][1]} k-evaluations)",
                   "judge_error_bound": abs(hist[-2][0] - ref) / abs(ref),
                   "uniform_levels": levels, "uniform_k_evals_needed": uni_need,
                   "learned": {"k_evals": lr["n_evals"], "rel_err": rel(lr["integral"]),
                               "flagged": lr["flagged"], "seconds": t_learned,
                               "k_per_s": lr["n_evals"] / t_learned},
                   "probe_without_network": {"k_evals": pr["n_evals"],
                                             "rel_err": rel(pr["integral"])},
                   "uniform_seconds": t_uniform},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def kkr_bench(args):
    """k-loop with structure constants from the layer producer (no CPU leg:
    the oracle's per-k gstore takes ~0.3 s)."""
    from skkr_budapest_amd import gpu
    from skkr_budapest_amd import kkrsystems as K
    E = complex(0.5, 0.05)
    rel = 1 if args.kkr_rel else 0
    wl = K.nb110_workload(E, P=20, lmax=3, rel=1) if rel else K.nb110_workload(E)
    st = wl["structure"]
    k, w = K.reciprocal_mesh(st["a1"], st["a2"], 64)
    nk = len(w)
    lk = (2 if rel else 1) * (st["lmax"] + 1) ** 2
    print(f"bench: kkr Nb(110) lmax={st['lmax']} rel={rel} lk={lk} nk={nk} E={E}", file=sys.stderr, flush=True)
    with gpu.Structure(**st) as sc, gpu.Context(lk=lk, sites_per_pl=st["s"], n_pl=wl["P"]) as ctx:
        ctx.set_tinv(wl["tinv"], wl["tinv_lead"], wl["tinv_lead"])
        ctx.set_structure_kkr(sc, wl["alpha"], rel=rel)
        for i in range(args.warmup):
            ctx.tau_bz(E, k, w)
        times, stats = [], []
        for i in range(args.steps):
            t0 = time.perf_counter()
            r = ctx.tau_bz(E, k, w, timing_detail=True)
            times.append(time.perf_counter() - t0)
            stats.append(r["stats"])
            print(f"bench: step {i} {times[-1] * 1e3:.1f} ms {r['stats']}", file=sys.stderr, flush=True)
    ms = 1e3 * float(np.mean(times))
    stl = stats[-1]
    out = {
        "metric": METRIC, "value": nk / (ms * 1e-3), "unit": "k-points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic t^-1 (square wells), reference-form structure constants",
        "config": {"workload": f"Nb(110) stack on the layer producer: lmax={st['lmax']} "
                               f"{'kappa-mu' if rel else 'lm'} lk={lk} s={st['s']} nextra={st['nextra']} "
                               f"P={wl['P']} n={st['s'] * lk} E={E} 64x64 k",
                   "precision": "fp64", "nk": nk, "mean_iter_l": stl["mean_iter_l"],
                   "mean_iter_r": stl["mean_iter_r"],
                   "phase_ms": {"structure_constants": stl["ms_fourier"], "decimation": stl["ms_decimation"],
                                "sweeps": stl["ms_sweeps"], "accumulate": stl["ms_accumulate"]}},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def kkr_sample_bench(args):
    """The energy loop of the reference's sample 02-layer-scf on the layer
    producer: the Nb(110) C2vx stack (input_geo.in: NPRC 3 x 4 layers, a =
    4.667470441, beta = 0.7071), lmax 2 relativistic (kappa-mu, lk 18), the
    16 contour energies of input_rsp.in:32 (zmesh) with the k-sets of
    input_rsp.in:35-39 (intbz 1 wedges of kmesh2d: 253 ... 1953 k) and the C2v
    group sum of the reference's rsp sphbas (the matrices from pointgroup).  t^-1 and screening are square
    wells (the sample's potentials are not used)."""
    from skkr_budapest_amd import contour, gpu
    from skkr_budapest_amd import kkrsystems as K
    from skkr_budapest_amd import kmesh as KM
    z, _ = contour.semicircle(16, -1.0, -0.0183132794260, 150.0)
    ksets = [21] * 10 + [31] * 2 + [41] * 2 + [51] + [61]
    st = K.nb110_workload(z[0], P=3, lmax=2, rel=1)["structure"]
    lat = KM.primkr2d(2, "C2vx", st["a1"][0], 0.70710678120, 0.0)
    meshes = {s: KM.kmesh2d(1, s, lat) for s in sorted(set(ksets))}
    from skkr_budapest_amd import pointgroup as PG
    dg, dg1, _ = PG.group_matrices("C2v", lmax=2, rel=True)
    nk_total = sum(len(meshes[s][1]) for s in ksets)
    print(f"bench: sample 02-layer-scf energy loop, 16 energies, {nk_total} k-evaluations", file=sys.stderr,
          flush=True)

    def one_energy(ctx, sc, ie):
        E = complex(z[ie])
        wl = K.nb110_workload(E, P=3, lmax=2, rel=1)
        ctx.set_tinv(wl["tinv"], wl["tinv_lead"], wl["tinv_lead"])
        ctx.set_structure_kkr(sc, wl["alpha"], rel=1)
        k, w = meshes[ksets[ie]]
        t0 = time.perf_counter()
        r = ctx.tau_bz(E, k, w, raise_on_error=False)
        if r["rc"] not in (0, 6):  # 6: decimation reached itermax somewhere (surfgfd warns and goes on)
            raise gpu.SkkrError(r["rc"], f"tau_bz at E={E}")
        return time.perf_counter() - t0, r

    with gpu.Structure(**st) as sc, gpu.Context(lk=18, sites_per_pl=st["s"], n_pl=3) as ctx:
        ctx.set_point_group(dg, dg1)
        for _ in range(args.warmup):
            one_energy(ctx, sc, 0)
            one_energy(ctx, sc, 15)
        loops, per_e = [], []
        for step in range(args.steps):
            tot, rows = 0.0, []
            for ie in range(16):
                dt, r = one_energy(ctx, sc, ie)
                tot += dt
                rows.append({"E": [z[ie].real, z[ie].imag], "nk": len(meshes[ksets[ie]][1]),
                             "ms": 1e3 * dt, "mean_iter": 0.5 * float(np.mean(r["iters"])),
                             "max_iter": int(r["iters"].max()), "k_at_itermax": int(np.sum(r["status"] != 0))})
            loops.append(tot)
            per_e = rows
            print(f"bench: contour {step}: {tot:.3f} s", file=sys.stderr, flush=True)
        ctx.set_structure_kkr(None, None)
    sec = float(np.mean(loops))
    out = {
        "metric": METRIC, "value": nk_total / sec, "unit": "k-points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * sec, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "sample 02-layer-scf geometry, contour and k-sets; square-well t^-1 and screening",
        "config": {"workload": "sample 02-layer-scf energy loop: Nb(110) C2vx, 3 PLs x 4 layers, lmax 2 "
                               "kappa-mu (lk 18, n 72), 16 contour energies, intbz 1 wedges with the C2v "
                               "group sum",
                   "precision": "fp64", "k_evaluations": nk_total, "per_energy": per_e},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true",
                    help="time only the FP64 k-loop (default: FP64, then the INT8 and split-BF16 mixed modes)")
    ap.add_argument("--modes", default="",
                    help="comma list of k-loop modes to time (fp64, mixed_int8, mixed_bf16); "
                         "default all (fp64 alone with --fp64-only)")
    ap.add_argument("--batch", type=int, default=0,
                    help="k-points per launch group (0: the library's choice)")
    ap.add_argument("--kkr", action="store_true",
                    help="k-loop on the layer structure-constant producer (Nb(110) stack)")
    ap.add_argument("--kkr-sample", action="store_true",
                    help="the energy loop of the reference's sample 02-layer-scf on the layer producer")
    ap.add_argument("--kkr-rel", action="store_true",
                    help="with --kkr: the relativistic lmax-3 stack of config 1's shape "
                         "(lk 32, 4 sites/PL, 20 PLs) on the producer")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the k-points a multi-GPU run would shard are all taken by one device.",
              file=sys.stderr)
        sys.exit(2)

    from skkr_budapest_amd import gpu
    from skkr_budapest_amd import systems as S

    if args.kkr_sample:
        return kkr_sample_bench(args)
    if args.kkr:
        return kkr_bench(args)
    cfg = S.CONFIGS[args.config]
    sy = S.make_system(cfg)
    if args.config in (2, 3):
        return contour_bench(args, cfg, sy)
    if args.config == 4:
        return mesh_bench(args, cfg, sy)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    nk = len(w)
    print(f"bench: config {args.config} n={cfg.n} P={cfg.P} nk={nk} E={E}", file=sys.stderr,
          flush=True)

    modes = [("fp64", gpu.FP64)] + ([] if args.fp64_only else [("mixed_int8", gpu.MIXED_INT8),
                                                                ("mixed_bf16", gpu.MIXED_BF16)])
    if args.modes:
        want = args.modes.split(",")
        modes = [md for md in modes if md[0] in want]
    # the modes' timed steps alternate (fp64, int8, bf16, fp64, ...) so that
    # each sees the same clock and temperature of the device
    ctxs = {}
    for name, prec in modes:
        ctxs[name] = gpu.Context.for_system(sy, precision=prec, batch=args.batch)
        ctxs[name].set_tinv(*sy.tinv(E))
    runs = kloop_steps(ctxs, E, k, w, args)
    for ctx in ctxs.values():
        ctx.close()
    head_name = "mixed_int8" if "mixed_int8" in runs else list(runs)[0]
    head = runs[head_name]
    value = nk / (head["ms"] * 1e-3)
    cfg_out = {"workload": f"cfg{args.config}: lk={cfg.lk} s={cfg.s} P={cfg.P} "
                           f"n={cfg.n} N_R={cfg.n_R} E={E} {cfg.mesh}x{cfg.mesh} k",
               "precision": head_name, "nk": nk}
    for name, r in runs.items():
        cfg_out[name] = mode_summary(r, nk, cfg)
    from skkr_budapest_amd import observables as O
    if "mixed_int8" in runs and "fp64" in runs:
        cfg_out["mixed_over_fp64"] = runs["fp64"]["ms"] / runs["mixed_int8"]["ms"]
        cfg_out["mixed_error_vs_fp64"] = O.error_report(runs["mixed_int8"]["tau_bz"], runs["fp64"]["tau_bz"])
        # the library's defaults for SKKR_MIXED_INT8 (skkr_create)
        cfg_out["mixed_int8_digits"] = ("decimation 6 (5 from Im E = 0.1, FP64 below fp64_dec_below_imE; "
                                        "4 in a chain's iterations after |alpha'| fell to 2^-6 of its first "
                                        "value, for n >= 64; eps^-1 of the iteration before kept from 2^-20 on), "
                                        "sweeps 5 (6 below fp64_dec_below_imE)")
    if "mixed_bf16" in runs and "fp64" in runs:
        cfg_out["mixed_bf16_over_fp64"] = runs["fp64"]["ms"] / runs["mixed_bf16"]["ms"]
        cfg_out["mixed_bf16_error_vs_fp64"] = O.error_report(runs["mixed_bf16"]["tau_bz"],
                                                             runs["fp64"]["tau_bz"])
    rl = rooflines(head)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "k-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded G_r(R), t^-1)",
        "config": cfg_out,
        "roofline": rl[0],
        "rooflines": rl,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sy, E, k, w)
    print(json.dumps(out))


def kloop_steps(ctxs, E, k, w, args):
    """Timed steps on the default path (decimation rounds and sweeps replayed
    from HIP graphs), the modes taking turns step by step; then one step per
    mode with eager launches, each bracketed by HIP events, for the per-kernel
    figures (the replays hide their launches)."""
    for name, ctx in ctxs.items():
        for i in range(args.warmup):
            r = ctx.tau_bz(E, k, w)
            print(f"bench: {name} warmup {i} {r['stats']['ms_total']:.1f} ms", file=sys.stderr, flush=True)
    times = {name: [] for name in ctxs}
    stats = {name: [] for name in ctxs}
    last = {}
    for i in range(args.steps):
        for name, ctx in ctxs.items():
            t0 = time.perf_counter()
            r = ctx.tau_bz(E, k, w)
            times[name].append(time.perf_counter() - t0)
            stats[name].append(r["stats"])
            last[name] = r
            print(f"bench: {name} step {i} {times[name][-1] * 1e3:.1f} ms {r['stats']}", file=sys.stderr,
                  flush=True)
    runs = {}
    for name, ctx in ctxs.items():
        ctx.set_option("graphs", 0)
        d = ctx.tau_bz(E, k, w, timing_detail=True)
        ctx.set_option("graphs", 1)
        print(f"bench: {name} detail step {d['stats']['ms_total']:.1f} ms {d['stats']}", file=sys.stderr,
              flush=True)
        runs[name] = {"ms": 1e3 * float(np.mean(times[name])), "stats": stats[name], "detail": [d["stats"]],
                      "tau_bz": last[name]["tau_bz"], "ms_eager": d["stats"]["ms_total"]}
    return runs


def mode_summary(r, nk, cfg):
    st = r["stats"][-1]
    it_l, it_r = st["mean_iter_l"], st["mean_iter_r"]
    return {"k_points_per_s": nk / (r["ms"] * 1e-3), "ms_per_step": r["ms"],
            "ms_per_step_eager_timed": r["ms_eager"],
            "mean_iter_l": it_l, "mean_iter_r": it_r, "n_launches": st["n_launches"],
            "n_launches_eager": r["detail"][-1]["n_launches"],
            "algorithmic_tflops": nk * algorithmic_flops_per_k(cfg.n, cfg.P, it_l, it_r)
            / (r["ms"] * 1e-3) / 1e12,
            "phase_ms": {p: st[f"ms_{p}"] for p in ("fourier", "decimation", "sweeps", "accumulate")}}


def rooflines(run):
    """Roofline entries of the GEMM and inversion kernels of the run, the one
    with more device time first.  achieved = algorithmic flops (active
    decimation chains only) / summed HIP-event durations of its launches (the
    eager step of the run)."""
    stats = run["detail"]
    out = []
    for cls in ("dom", "inv"):
        launches = sum(s[f"{cls}_launches"] for s in stats)
        ms = sum(s[f"{cls}_ms"] for s in stats)
        flops = sum(s[f"{cls}_flops"] for s in stats)
        nbytes = sum(s[f"{cls}_bytes"] for s in stats)
        if not launches or ms <= 0:
            continue
        alg = flops / (ms * 1e-3) / 1e12
        per = flops / launches
        if cls == "dom" and stats[-1]["dom_name"].startswith("zgemm_i8"):
            exe = sum(s["dom_flops_exec"] for s in stats) / (ms * 1e-3) / 1e12
            ent = {"bound": "mfma", "kernel": stats[-1]["dom_name"], "achieved": exe,
                   "peak": PEAK_INT8_TOPS, "unit": "TOP/s", "frac": exe / PEAK_INT8_TOPS,
                   "achieved_is": "executed INT8 MFMA ops",
                   "algorithmic_tflops_fp64_equivalent": alg,
                   "algorithmic_frac_of_fp64_peak": alg / PEAK_FP64_TFLOPS}
        elif cls == "dom" and stats[-1]["dom_name"].startswith("zgemm_split"):
            exe = sum(s["dom_flops_exec"] for s in stats) / (ms * 1e-3) / 1e12
            ent = {"bound": "mfma", "kernel": "zgemm_split_kernel", "achieved": exe,
                   "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "frac": exe / PEAK_BF16_TFLOPS,
                   "achieved_is": "executed BF16 MFMA flops",
                   "algorithmic_tflops_fp64_equivalent": alg,
                   "algorithmic_frac_of_fp64_peak": alg / PEAK_FP64_TFLOPS}
        else:
            name = stats[-1]["dom_name"] if cls == "dom" else "gjinv_blk_f64_kernel"
            ent = {"bound": "mfma", "kernel": name, "achieved": alg, "peak": PEAK_FP64_TFLOPS,
                   "unit": "TFLOP/s", "frac": alg / PEAK_FP64_TFLOPS}
        ent.update({"traffic": (nbytes / launches) * (GEMM_TRAFFIC_PER_ALGO_BYTE if cls == "dom" else 1.0),
                    "traffic_unit": "algorithmic bytes per launch x PMC calibration" if cls == "dom"
                    else "algorithmic bytes per launch",
                    "algorithmic_flops_per_launch": per, "avg_launch_us": 1e3 * ms / launches,
                    "launches": launches, "device_ms": ms})
        out.append(ent)
    out.sort(key=lambda e: -e["device_ms"])
    return out


if __name__ == "__main__":
    main()
