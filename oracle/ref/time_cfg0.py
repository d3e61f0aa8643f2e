"""Times the reference's own k-loop subroutines on config 0 (and on a
larger n = 64 stack) beside the C oracle, on this machine's cores, for the
CPU baseline of the benchmark.

Config 0 as the reference runs it: two nonrelativistic spin channels of
9 x 9 site blocks (skkr_kpar_v6/nrsp-scf/skkr.f:258), 2 sites per principal
layer, 8 layers, E = 0.5 + 0.05i, the full 16 x 16 mesh.  The reference path
is oracle/_ref/refdrv_kpar (the kpar dugo, surfgfd, packer, tau2d, gjinv and
matops compiled with flang -O2 by oracle/ref/Makefile, the level of
skkr_kpar_v6/makefile:27), run as NPROC single-threaded processes that take
the k-points round-robin, modulo(ik-1, nprocs) == myrank
(nrsp-scf/cpacoord.f:147-149), with their partial BZ sums added (the
reference's MPI_Allreduce).  The C oracle runs the same two channels with
NPROC threads splitting k round-robin.  Both results are checked against each
other.  Prints one JSON line: k-points/s per energy of both (both spin
channels of every k) and their ratio, the cross-normalisation factor.

The second workload ("n64") is the largest shape the reference's unchanged
param.h admits (skkr_dbogo_bsf: dense site blocks <= 18, <= 4 sites per
layer, <= 4 layers): lk = 16, 4 sites, 4 layers, the Hermitian t^-1 law of
the relativistic configurations, E = 0.6 + 0.02i, 16 x 16 k, through the
dbogo twins (refdrv_dbogo), where process start-up is negligible.

Usage (needs /root/reference, flang):  python oracle/ref/time_cfg0.py [nproc] [repeats]
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from skkr_budapest_amd import systems as S  # noqa: E402

REFDIR = os.path.join(ROOT, "oracle", "_ref")


def write_input(path, sy, E, k, w, idx):
    """refdrv input (refdrv.F header) for the k-points idx of the mesh."""
    cfg = sy.cfg
    tint, tl, tr = sy.tinv(E)
    with open(path, "wb") as f:
        np.array([cfg.lk, cfg.s, cfg.P, len(idx), cfg.itermax, 0], np.int32).tofile(f)
        np.array([cfg.errmax], np.float64).tofile(f)
        for t in (tint, tl, tr):
            S.colmajor(t).tofile(f)
        w[idx].astype(np.float64).tofile(f)
        for ik in idx:
            ph = np.exp(-1j * (sy.R @ k[ik]))
            S.colmajor(np.tensordot(ph, sy.G_intra, axes=(0, 0))).tofile(f)
            S.colmajor(np.tensordot(ph, sy.G_inter, axes=(0, 0))).tofile(f)


def time_workload(fam, chans, E, k, w, nproc, reps):
    """Wall time of the reference (nproc processes, round-robin k, every
    channel in turn) and of the oracle (nproc threads), median of reps."""
    nk = len(w)
    exe = os.path.join(REFDIR, f"refdrv_{fam}")
    ref_t, port_t = [], []
    with tempfile.TemporaryDirectory() as td:
        files = []
        for r in range(nproc):
            idx = np.array([ik for ik in range(nk) if ik % nproc == r])
            per = []
            for sp, sy in enumerate(chans):
                fin = os.path.join(td, f"in_{r}_{sp}.bin")
                write_input(fin, sy, E, k, w, idx)
                per.append((fin, os.path.join(td, f"out_{r}_{sp}.bin")))
            files.append(per)
        for _ in range(reps):
            t0 = time.perf_counter()
            procs = []
            for per in files:
                cmd = "ulimit -s unlimited; " + "; ".join(
                    f'"{exe}" "{fin}" "{fout}" > /dev/null' for fin, fout in per)
                procs.append(subprocess.Popen(["bash", "-c", cmd]))
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("refdrv failed")
            ref_t.append(time.perf_counter() - t0)
        # the ranks' partial sums added (the reference's MPI_Allreduce)
        ns, lk = chans[0].cfg.P * chans[0].cfg.s, chans[0].cfg.lk
        ref_bz = [np.zeros((ns, lk, lk), complex) for _ in chans]
        for per in files:
            for sp, (_, fout) in enumerate(per):
                raw = np.fromfile(fout, np.complex128)
                ref_bz[sp] += S.from_colmajor(raw[:ns * lk * lk].reshape(ns, lk, lk))
    for _ in range(reps):
        t0 = time.perf_counter()
        outs = [oracle.tau_bz(sy, E, k, w, nworkers=nproc) for sy in chans]
        port_t.append(time.perf_counter() - t0)
    err = max(np.linalg.norm(o["tau_bz"] - r) / np.linalg.norm(r) for o, r in zip(outs, ref_bz))
    ref_rate = nk / float(np.median(ref_t))
    port_rate = nk / float(np.median(port_t))
    return {"reference_k_per_s": ref_rate, "port_k_per_s": port_rate,
            "reference_over_port": ref_rate / port_rate, "reference_seconds": ref_t,
            "port_seconds": port_t, "tau_bz_rel_diff": err}


def main():
    nproc = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    out = {"cores": nproc,
           "cpu": open("/proc/cpuinfo").read().split("model name")[1].split("\n")[0].strip(": \t")}
    c0 = S.make_system(S.CONFIGS[0])
    k, w = S.uniform_kmesh(S.CONFIGS[0].mesh)
    r = time_workload("kpar", [S.spin_channel(c0, sp) for sp in (0, 1)], S.CONFIGS[0].energies[0], k, w,
                      nproc, reps)
    r["workload"] = ("config 0 as the reference runs it: two spin channels lk=9, s=2, P=8, 13 R, "
                     "E=0.5+0.05i, 16x16 k")
    out["cfg0"] = r
    c64 = S.Config("time-n64", 73, lk=16, s=4, P=4, r2max=8, tmodel="herm", energies=(0.6 + 0.02j,), mesh=16)
    k, w = S.uniform_kmesh(c64.mesh)
    r = time_workload("dbogo", [S.make_system(c64)], c64.energies[0], k, w, nproc, reps)
    r["workload"] = "n64: lk=16, s=4, P=4, 25 R, Hermitian t^-1 law, E=0.6+0.02i, 16x16 k (dbogo twins)"
    out["n64"] = r
    print(json.dumps(out))


if __name__ == "__main__":
    main()
