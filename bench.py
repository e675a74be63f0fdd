"""Benchmark: Sobol points/s into H and time to the lowest 10 levels.

Workload (default): water hexamer relaxed under q-TIP4P/F (6 H2O, 48 modes),
basis of total quanta <= 2 (1225 states), 2^20 Sobol points (skip 2^30) in
4096-point tiles.  One step = one full pass of the hot path over the 2^20
points (sample -> map -> potential -> basis -> SYRK into the FP64 running H)
followed by the lowest 10 levels on the GPU.  Both GPU forms are run: the
plain FP64 form and the mixed form (FP32 potential and basis, 3-term
split-BF16 MFMA SYRK folded into FP64 H).  `value` is the mixed form's
points/s into H; the FP64 form, time-to-levels of both and the level
difference are reported beside it.  The CPU baseline is the FP64 oracle path
(Sobol/BSM, map, serial potential loop, get_bra, scipy/OpenBLAS dsyr2k) on a
bounded prefix of the same points.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--points LOG2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=20, help="log2 of Sobol points per step")
    ap.add_argument("--config", default="hexamer_qtip4pf")
    ap.add_argument("--nmodes2", type=int, default=48)
    ap.add_argument("--nmodes3", type=int, default=0)
    ap.add_argument("--levels", type=int, default=10)
    ap.add_argument("--cpu-sample", type=int, default=13, help="log2 points of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed")
    return ap.parse_args()


def run_form(c, sysd, mode, P, skip, steps, warmup, k):
    import scp3_amd
    c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
    c.set_basis(sysd["modes"], sysd["nmodes2"], sysd["nmodes3"])
    c.set_pes("qtip4pf" if "qtip4pf" in sysd["name"] else "ttm3f")
    c.set_precision(mode)
    c.set_tile(4096)
    for _ in range(warmup):
        c.reset()
        c.accumulate(skip, P)
        c.lowest_levels(k)
    c.sync()
    c.set_timing(True)
    t_acc = 0.0
    t_lev = 0.0
    ev = None
    for _ in range(steps):
        c.sync()
        t0 = time.perf_counter()
        c.reset()
        c.accumulate(skip, P)
        c.sync()
        t1 = time.perf_counter()
        ev = c.lowest_levels(k)
        c.sync()
        t2 = time.perf_counter()
        t_acc += t1 - t0
        t_lev += t2 - t1
    kt = {}
    for name in ("sobol", "map", "pes", "weight", "basis", "syrk", "symv", "symfill"):
        ms, cnt = c.kernel_time(name)
        kt[name] = (ms, cnt)
    c.set_timing(False)
    return dict(t_acc=t_acc / steps, t_lev=t_lev / steps, levels=ev, kernels=kt,
                n=c.basis_size())


def cpu_baseline(sysd, skip, logS):
    """FP64 oracle path on the host: bounded prefix of the same points."""
    import scipy.linalg.blas as blas

    from oracle import oracle
    from scp3_amd import AUTOKCALPMOL, BOHR
    S = 1 << logS
    M0 = sysd["omega"].size
    modes = sysd["modes"]
    n = oracle.basis_size(modes.size, sysd["nmodes2"], sysd["nmodes3"])
    cores = len(os.sched_getaffinity(0))
    t = {}
    t0 = time.perf_counter()
    y = oracle.sobol_normals(M0, skip, S)
    t["sample"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = BOHR * (sysd["x0"][None, :] + y @ sysd["MUa"].T)
    t["map"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    V = np.array([oracle.qtip4pf(r[i]) for i in range(S)]) / AUTOKCALPMOL
    t["pes"] = time.perf_counter() - t0
    bV = V - 0.5 * (y * sysd["omega"] * y).sum(1)
    t0 = time.perf_counter()
    bra = oracle.get_bra(y[:, modes], modes.size, sysd["nmodes2"], sysd["nmodes3"])
    t["basis"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    Mm = np.zeros((n, n), order="F")
    A = np.asfortranarray(bra.T)
    Bk = np.asfortranarray(bra.T * bV[None, :])
    Mm = blas.dsyr2k(0.5, A, Bk, beta=1.0, c=Mm, trans=0, lower=0, overwrite_c=1)
    t["syrk"] = time.perf_counter() - t0
    total = sum(t.values())
    return dict(value=S / total, unit="points/s", cores=cores, kind="port",
                sample="%d Sobol points (skip 2^30) of the same workload; phases %s s; "
                       "FP64 oracle path: C Sobol/BSM, numpy map, serial potential loop, "
                       "get_bra, OpenBLAS dsyr2k (threads = %d)"
                       % (S, {k: round(v, 4) for k, v in t.items()}, cores),
                phase_s=t)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the multi-GPU sharding is described in DESIGN.md)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import scp3_amd
    from scp3_amd import system

    sysd = system.prepare(system.input_path(a.config), nmodes2=a.nmodes2, nmodes3=a.nmodes3)
    sysd["name"] = a.config
    P = 1 << a.points
    skip = 1 << 30
    forms = a.forms.split(",")
    res = {}
    with scp3_amd.Context(0) as c:
        for f in forms:
            mode = scp3_amd.FP64 if f == "fp64" else scp3_amd.MIXED
            print("bench: form %s, %d points, n=%s" % (f, P, "?"), file=sys.stderr, flush=True)
            res[f] = run_form(c, sysd, mode, P, skip, a.steps, a.warmup, a.levels)
    n = res[forms[0]]["n"]
    main_form = "mixed" if "mixed" in res else forms[0]
    r = res[main_form]
    # dominant kernel: the SYRK; algorithmic flops per launch = n(n+1) * tile points
    ms, cnt = r["kernels"]["syrk"]
    avg_s = ms / 1e3 / max(cnt, 1)
    tile = 4096
    flops = float(n) * (n + 1) * tile
    peak = PEAK_BF16_TFLOPS if main_form == "mixed" else PEAK_FP64_TFLOPS
    achieved = flops / avg_s / 1e12 if avg_s > 0 else 0.0
    out = {
        "metric": "Sobol pts/s into H + time to lowest-10 levels at 0.5 cm^-1, GPU forms vs plain",
        "value": P / r["t_acc"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * (r["t_acc"] + r["t_lev"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3/fp32 SYRK + fp64 H" if main_form == "mixed" else "f64",
        "data": "synthetic: seed-built, relaxed cluster; Sobol points",
        "config": {"workload": "%s n=%d, 2^%d Sobol points, 4096-point tiles, lowest %d levels"
                               % (a.config, n, a.points, a.levels),
                   "basis_states": n, "points": P, "form": main_form},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "syrk (%s)" % main_form, "avg_launch_ms": avg_s * 1e3,
                     "flops_per_launch": flops},
        "forms": {},
    }
    for f, rr in res.items():
        out["forms"][f] = {
            "points_per_s": P / rr["t_acc"],
            "time_to_levels_s": rr["t_acc"] + rr["t_lev"],
            "eigensolve_s": rr["t_lev"],
            "levels_cm1": [round(float(v) * scp3_amd.AUTOCM, 4) for v in rr["levels"]],
            "kernel_ms_per_step": {k: round(v[0] / a.steps, 4) for k, v in rr["kernels"].items() if v[1]},
        }
    if "fp64" in res and "mixed" in res:
        d = np.abs(res["mixed"]["levels"] - res["fp64"]["levels"]).max() * scp3_amd.AUTOCM
        out["mixed_vs_fp64_max_level_diff_cm1"] = float(d)
        out["speedup_time_to_levels_mixed_vs_fp64"] = (
            (res["fp64"]["t_acc"] + res["fp64"]["t_lev"]) / (r["t_acc"] + r["t_lev"]))
    if not a.no_cpu:
        cb = cpu_baseline(sysd, skip, a.cpu_sample)
        cb.pop("phase_s")
        out["cpu_baseline"] = cb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
