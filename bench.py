"""Benchmark: Sobol points/s into H and time to the lowest levels.

Default workload (preset c3): the water hexamer relaxed under TTM3-F (6
H2O, 48 modes), basis of total quanta <= 3 (20825 states), 2^21 scrambled
Sobol points (seed 0, skip 2^30) per step in 32768-point tiles (the
automatic tile).  One step = one full pass of the hot path over the 2^21
points (sample -> map -> TTM3-F -> weight -> basis -> SYRK into the FP64
running H) followed by the lowest 10 levels on the GPU (the K timed
accumulate passes run first, with per-kernel event timers, then K timed
eigensolves of the last pass's H without them).  Three GPU forms run: the
mixed form (FP32 potential and basis, 3-term split-BF16 MFMA SYRK folded
into FP64 H; `value` is its points/s into H), the plain FP64 form (at most
2 timed steps) and the MLP control variate (2^24 surrogate + 2^23
true-potential points, one timed step, after its training).  Beside them:
the level difference and H error budget of mixed against FP64 on the same
points, each form's time to the lowest 10 levels within 0.5 cm^-1 (this
run's rates x the points the committed convergence record needs), the
mixed SYRK's roofline, and the CPU baseline: the -G path as SCP3 runs it
(the reference's potential in a serial point loop, get_bra under OpenMP,
OpenBLAS dsyr2k) on a bounded prefix of the same points.

Other presets: c1 (dimer, 91 states, 2^16 points), c2 (hexamer under
q-TIP4P/F, 1225 states, 2^20 points, FP64 and mixed), c4 (c3 with the MLP
control variate as configured: 2^24 surrogate + 2^19 true points against
plain FP64 at 2^21), c5 (20-water cluster under TTM3-F, 174 modes, 15400
states, MLP 174-512-512-512-1 on 2^19 pairs, 2^24 surrogate + 2^20 true
points against plain FP64 at 2^20, lowest 20 levels); c3g (c3 with the
gradient estimator, Stein form, 2^20 points, FP64 against mixed); scp1 /
scp1q (SCP1's Gaussian averages: hexamer under TTM3-F at 2^18 points /
under q-TIP4P/F at 2^20 points, one step = one SCP1 iteration's averages).

Forms: fp64, mixed, cv (above) and grad / stein: the gradient (Stein-
identity) estimator, the reference's default without -G (grad: the
reference's element placement; stein: every element from the Stein
identity), in FP64, or with the rank-2 term on split-BF16 MFMA
(grad_mixed, stein_mixed), e.g. --preset c3 --forms stein,stein_mixed.
The JSON line's main form is mixed, else cv, else the last form listed.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--preset c3]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Davidson residual tolerance (hartree) for the timed levels: a Ritz value
# lies within the residual norm of an eigenvalue of the symmetric matrix, so
# 1e-6 Eh (0.22 cm^-1) meets the metric's 0.5 cm^-1 in the worst case
LEVEL_TOL = 1e-6

PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0

PRESETS = {
    "c1": dict(config="dimer_qtip4pf", n2=12, n3=0, points=16, tile=0, levels=5, forms="fp64,mixed"),
    "c2": dict(config="hexamer_qtip4pf", n2=48, n3=0, points=20, tile=0, levels=10, forms="fp64,mixed"),
    # the default: the plain FP64 and mixed forms over 2^21 points per step
    # (FP64 capped at 2 timed steps, 15 s each), and the control variate with
    # its correction term over the 2^23 true points that 0.5 cm^-1 needs
    # (2^24 surrogate points, one timed step)
    "c3": dict(config="hexamer_ttm3f", n2=48, n3=48, points=21, tile=0, levels=10, forms="mixed,fp64,cv",
               cv_nn=24, cv_t=23, mlp_hidden=256, mlp_steps=20000, mlp_batch=4096, mlp_pairs=18,
               caps={"fp64": (2, 1), "cv": (1, 1)}),
    "c4": dict(config="hexamer_ttm3f", n2=48, n3=48, points=21, tile=0, levels=10, forms="fp64,cv",
               cv_nn=24, cv_t=19, mlp_hidden=256, mlp_steps=20000, mlp_batch=4096, mlp_pairs=18),
    "c5": dict(config="w20_ttm3f", n2=174, n3=0, points=20, tile=0, levels=20, forms="fp64,cv",
               cv_nn=24, cv_t=20, mlp_hidden=512, mlp_steps=20000, mlp_batch=4096, mlp_pairs=19),
    # c3 with the gradient (Stein) estimator: FP64 against the mixed form
    "c3g": dict(config="hexamer_ttm3f", n2=48, n3=48, points=20, tile=0, levels=10, forms="stein,stein_mixed"),
    # SCP1's Gaussian averages (energy, gradient and the gradient-displacement
    # moments) over the Sobol points, one SCP1 iteration's worth
    "scp1": dict(kind="scp1", config="hexamer_ttm3f", pes="ttm3f", points=18),
    "scp1q": dict(kind="scp1", config="hexamer_qtip4pf", pes="qtip4pf", points=20),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--preset", default="c3", choices=sorted(PRESETS))
    ap.add_argument("--points", type=int, default=None, help="log2 of Sobol points per step")
    ap.add_argument("--forms", default=None)
    ap.add_argument("--tile", type=int, default=None, help="points per tile (0 = automatic)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline time budget")
    ap.add_argument("--cpu-chunk", type=int, default=1024,
                    help="points per CPU baseline block (the reference's -B, SCP3.cpp:58)")
    ap.add_argument("--converge-summary", nargs="+", default=None,
                    help="JSON-line outputs of --converge runs: write the convergence record of the preset")
    ap.add_argument("--merge-record", action="store_true",
                    help="with --converge-summary: keep the committed record's checkpoints and add the new ones")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cv-terms", type=int, default=1)
    ap.add_argument("--scramble", type=int, default=0, help="seed of the affine digit scramble of the Sobol points")
    ap.add_argument("--no-scramble", action="store_true", help="the reference's unscrambled Sobol sequence")
    ap.add_argument("--mixed-kernel", type=int, default=None, help="SYRK kernel variant (scp3g_set_mixed_kernel)")
    ap.add_argument("--converge", default=None, help="convergence study of one form (fp64, mixed, stein_mixed, cv)")
    ap.add_argument("--min-log2", type=int, default=21)
    ap.add_argument("--max-log2", type=int, default=24)
    ap.add_argument("--seeds", default="0", help="scramble seeds of the convergence study")
    ap.add_argument("--level-tol", type=float, default=1e-7, help="Davidson residual tolerance (hartree)")
    ap.add_argument("--cv-t", type=int, default=None, help="log2 of true-potential points of the control variate")
    ap.add_argument("--cv-nn", type=int, default=None, help="log2 of surrogate points of the control variate")
    ap.add_argument("--mlp-steps", type=int, default=None, help="Adam steps of the surrogate's training")
    ap.add_argument("--mlp-pairs", type=int, default=None, help="log2 of the surrogate's training pairs")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration(nbytes):
    """One plain device-to-device copy of nbytes (torch's copy kernel): the
    known-size reference that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "check": float(dst[-1].item())}))


GRAD_FORMS = ("grad", "stein", "grad_mixed", "stein_mixed")


def setup(c, sysd, pr, mode, form=None):
    import scp3_amd
    est = {"grad": scp3_amd.EST_GRADIENT, "stein": scp3_amd.EST_GRADIENT_STEIN}
    c.set_estimator(est.get((form or "").replace("_mixed", ""), scp3_amd.EST_ENERGY))
    c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
    c.set_scramble(pr.get("scramble"))
    c.set_basis(sysd["modes"], sysd["nmodes2"], sysd["nmodes3"])
    c.set_pes("ttm3f" if "ttm3f" in pr["config"] else "qtip4pf")
    c.set_precision(mode)
    c.set_tile(pr["tile"])
    if pr.get("mixed_kernel") is not None:
        c.set_mixed_kernel(pr["mixed_kernel"])


def converge(a, pr, sysd):
    """Convergence study of one form: levels after a growing number of Sobol
    points (one running sum, levels read at each checkpoint; cv: the
    correction term over 2^cv_t true-potential points once, the surrogate
    term grown over the checkpoints), one JSON line per checkpoint with the
    accumulate wall time up to it."""
    import scp3_amd
    form = a.converge
    mode = scp3_amd.FP64 if form in ("fp64", "grad", "stein") else scp3_amd.MIXED
    k = pr["levels"]
    skip = 1 << 30
    cps = []
    for m in range(a.min_log2, a.max_log2 + 1):
        cps.append(1 << m)
        if m < a.max_log2:
            cps.append(3 << (m - 1))
    with scp3_amd.Context(0) as c:
        c.L.scp3g_set_cv_terms(c.handle, a.cv_terms)
        for seed in [int(s) for s in a.seeds.split(",")]:
            pr["scramble"] = seed
            setup(c, sysd, pr, mode, form)
            train_s = 0.0
            if form == "cv":
                c.sync()
                t0 = time.perf_counter()
                c.mlp_init(pr["mlp_hidden"], seed=0)
                c.mlp_sample_data(1 << 40, 1 << pr["mlp_pairs"])
                loss = c.mlp_train(pr["mlp_steps"], pr["mlp_batch"], 1e-3)
                c.sync()
                train_s = time.perf_counter() - t0
                print(json.dumps({"form": form, "seed": seed, "train_s": train_s,
                                  "final_loss": float(loss[-100:].mean())}), flush=True)
            c.reset()
            done, t_acc, Ct = 0, 0.0, None
            if form == "cv":
                # the correction term once over 2^cv_t true-potential points
                # (the first of the surrogate's seeds), then the surrogate
                # term grown checkpoint by checkpoint in its own sum:
                # H = S / P_nn + C / P_t
                Pt = 1 << pr["cv_t"]
                c.sync()
                t0 = time.perf_counter()
                c.accumulate_correction(skip, Pt)
                c.sync()
                t_acc = time.perf_counter() - t0
                Ct = c.h_download(raw=True) / Pt
                c.reset()
            for P in cps:
                # pieces of 2^25 points (2^23 in the slower FP64 forms), each
                # with a progress line on stderr
                piece = 1 << (23 if mode == scp3_amd.FP64 else 25)
                while done < P:
                    m = min(piece, P - done)
                    c.sync()
                    t0 = time.perf_counter()
                    if form == "cv":
                        c.accumulate_surrogate(skip + done, m)
                    else:
                        c.accumulate(skip + done, m)
                    c.sync()
                    t_acc += time.perf_counter() - t0
                    done += m
                    print("converge: %s seed %d, %d points, %.1f s" % (form, seed, done, t_acc),
                          file=sys.stderr, flush=True)
                t1 = time.perf_counter()
                if form == "cv":
                    M = c.h_download()
                    M += Ct
                    ev = c.matrix_lowest_levels(M, k, a.level_tol)
                    del M
                else:
                    ev = c.lowest_levels(k, a.level_tol)
                t_lev = time.perf_counter() - t1
                print(json.dumps({"form": form, "seed": seed, "points": P,
                                  "points_t": (1 << pr["cv_t"]) if form == "cv" else None,
                                  "accumulate_s": round(t_acc, 4), "train_s": round(train_s, 4),
                                  "eigensolve_s": round(t_lev, 4),
                                  "levels_cm1": [round(float(v) * scp3_amd.AUTOCM, 4) for v in ev]}), flush=True)


def run_form(c, sysd, pr, form, P, skip, steps, warmup, k, extra, keep_h=False):
    import scp3_amd
    mode = scp3_amd.FP64 if form in ("fp64", "grad", "stein") else scp3_amd.MIXED
    setup(c, sysd, pr, mode, form)
    train_s = 0.0
    if form == "cv":
        t0 = time.perf_counter()
        c.mlp_init(pr["mlp_hidden"], seed=0)
        c.mlp_sample_data(1 << 40, 1 << pr["mlp_pairs"])
        loss = c.mlp_train(pr["mlp_steps"], pr["mlp_batch"], 1e-3)
        c.sync()
        train_s = time.perf_counter() - t0
        extra["mlp_final_loss"] = float(loss[-100:].mean())
        extra["mlp_train_s"] = train_s

    def one():
        c.reset()
        if form == "cv":
            # correction points: the first 2^cv_t of the surrogate's Sobol sequence
            c.accumulate_cv(skip, 1 << pr["cv_nn"], skip, 1 << pr["cv_t"])
        else:
            c.accumulate(skip, P)

    for _ in range(warmup):
        one()
        c.lowest_levels(k, LEVEL_TOL)
    c.sync()
    # accumulate steps with the per-kernel event timers on (the roofline
    # needs the SYRK kernel's own time), then the eigensolves on the last
    # step's H with the timers off (their events would add host latency to
    # a latency-bound solver)
    c.set_timing(True)
    t_acc = t_lev = 0.0
    ev = None
    for _ in range(steps):
        c.sync()
        t0 = time.perf_counter()
        one()
        c.sync()
        t_acc += time.perf_counter() - t0
    kt = {}
    for name in ("sobol", "map", "pes", "weight", "mlp", "basis", "syrk", "syrk_mma", "grad_blocks"):
        kt[name] = c.kernel_time(name)
    c.set_timing(False)
    for _ in range(steps):
        c.sync()
        t1 = time.perf_counter()
        ev = c.lowest_levels(k, LEVEL_TOL)
        c.sync()
        t_lev += time.perf_counter() - t1
    npts = c.points_done()
    return dict(t_acc=t_acc / steps, t_lev=t_lev / steps, levels=ev, kernels=kt, n=c.basis_size(),
                points=npts, train_s=train_s, Mout=c.h_download() if keep_h else None)


def error_budget(Ma, Mb, la, lb):
    """Max-abs and Frobenius norm of dH = Ma - Mb over the upper triangles
    (absolute in hartree and relative to Mb's), and each level's difference."""
    import scp3_amd
    out = {}
    # column blocks keep the temporaries small at 20825 states
    mx = fro2 = ref_mx = ref_fro2 = 0.0
    n = Ma.shape[0]
    for c0 in range(0, n, 2048):
        A = np.triu(Ma[:, c0:c0 + 2048], -c0)
        B = np.triu(Mb[:, c0:c0 + 2048], -c0)
        D = A - B
        mx = max(mx, float(np.abs(D).max()))
        fro2 += float((D * D).sum())
        ref_mx = max(ref_mx, float(np.abs(B).max()))
        ref_fro2 += float((B * B).sum())
    out["dH_max_abs_hartree"] = mx
    out["dH_max_abs_rel"] = mx / ref_mx
    out["dH_frobenius_hartree"] = float(np.sqrt(fro2))
    out["dH_frobenius_rel"] = float(np.sqrt(fro2 / ref_fro2))
    out["level_diff_cm1"] = [round(float(v), 5) for v in (np.asarray(la) - np.asarray(lb)) * scp3_amd.AUTOCM]
    return out


def pmc_traffic(kernel, config, n, points_per_launch):
    """HBM bytes per launch of `kernel` on this workload from the committed
    FETCH_SIZE / WRITE_SIZE summary (rocprofv3 --pmc runs, scaled by a
    plain-copy calibration), or None when that workload was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic_syrk_x3_hexamer.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 6 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel and f[2] == config \
                and int(f[3]) == n and int(f[4]) == int(round(points_per_launch)):
            return float(f[5])
    return None


def cpu_baseline(sysd, pes, skip, scramble, bw, target_s, P_total):
    """The -G path as SCP3 runs it on the host, on the same scrambled points:
    per block of bw points (the reference's -B, SCP3.cpp:58) the sampler
    (oracle C Sobol + BSM), the serial potential loop of SCP3.cpp:205-225
    with the reference's own potential (oracle/_ref, else the oracle's
    restatement), get_bra under OpenMP over the points
    (ScaledMatrixElements.cpp:956-961) and dsyr2k('U', 'N', 0.5, bra,
    bra diag(bV), 1, M) in OpenBLAS (scipy's).  Blocks run until target_s;
    the rate is extrapolated linearly to P_total points."""
    import scipy.linalg.blas as blas

    from oracle import oracle
    modes = sysd["modes"]
    M0 = sysd["omega"].size
    n = oracle.basis_size(modes.size, sysd["nmodes2"], sysd["nmodes3"])
    fnptr, kind = oracle.pes_pointer(pes)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    t = dict(sample=0.0, pes=0.0, basis=0.0, syr2k=0.0)
    M = np.zeros((n, n), order="F")
    bra = np.zeros((bw, n))
    done = 0
    while done < P_total and sum(t.values()) < target_s:
        t0 = time.perf_counter()
        y = oracle.sobol_normals(M0, skip + done, bw, scramble=scramble)
        t1 = time.perf_counter()
        bV = oracle.block_weights(sysd, fnptr, y)
        t2 = time.perf_counter()
        oracle.block_bra(sysd, y, out=bra)
        t3 = time.perf_counter()
        ket = bra * bV[:, None]
        M = blas.dsyr2k(0.5, bra.T, ket.T, beta=1.0, c=M, trans=0, lower=0, overwrite_c=1)
        t4 = time.perf_counter()
        for key, dt in zip(t, (t1 - t0, t2 - t1, t3 - t2, t4 - t3)):
            t[key] += dt
        done += bw
    total = sum(t.values())
    rate = done / total
    return dict(value=rate, unit="points/s", cores=threads, kind="port",
                sample="%d scrambled Sobol points (seed %s, skip 2^30) of the same workload in %d-point blocks "
                       "(%.1f s): sampler %.2f s, serial %s potential loop (%s build) %.2f s, get_bra %.2f s, "
                       "OpenBLAS dsyr2k %.2f s; %d threads for get_bra and dsyr2k; extrapolated linearly "
                       "to %d points: %.0f s" % (done, scramble, bw, total, t["sample"], pes, kind, t["pes"],
                                                 t["basis"], t["syr2k"], threads, P_total, P_total / rate))


def scp1_modes(path):
    """q (bohr), mass, d and isqrtM_U of an input's Hessian at T = 0, as the
    SCP1 driver starts (SCP1.cpp:105-110; canonical eigenvector signs)."""
    from scp3_amd import AUTOCM, BOHR, system
    species, x, H = system.read_vladimir(path)
    mass, xs, Hs, _ = system.ohhohh(species, x, H)
    w, U = np.linalg.eigh(Hs, UPLO="U")
    U = U * np.sign(U[np.argmax(np.abs(U), axis=0), np.arange(U.shape[1])])[None, :]
    om = np.sqrt(np.abs(w))
    with np.errstate(divide="ignore"):
        d = np.where(om < 50.0 / AUTOCM, 0.0, 0.5 / om)
    return xs / BOHR, mass, d, U / np.sqrt(mass)[:, None]


def bench_scp1(a, pr):
    """One step = scp3g_scp1_average over 2^points Sobol points: normals,
    map, energy and gradient, moments, as one SCP1 iteration does."""
    import scp3_amd
    from scp3_amd import system
    P = 1 << pr["points"]
    skip = P * 2  # SCP1.h:56
    q, mass, d, iU = scp1_modes(system.input_path(pr["config"]))
    with scp3_amd.Context(0) as c:
        for _ in range(a.warmup):
            c.scp1_average(pr["pes"], q, d, iU, skip, P)
        c.set_timing(True)
        c.sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            U0, UX, UXq = c.scp1_average(pr["pes"], q, d, iU, skip, P)
        c.sync()
        t = (time.perf_counter() - t0) / a.steps
        kern = {k: c.kernel_time(k) for k in ("scp1_sobol", "scp1_map", "scp1_pes", "scp1_moments")}
    n3 = q.size
    out = {
        "metric": "SCP1 Gaussian averages: Sobol points/s (energy, gradient, <UX dq^T>)",
        "value": P / t,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seed-built relaxed cluster (scp3_amd/data), Sobol points",
        "config": {"workload": "%s: %s SCP1 averages, %d Cartesian dims, 2^%d Sobol points (skip 2^%d), "
                               "131072-point tiles" % (a.preset, pr["config"], n3, pr["points"], pr["points"] + 1),
                   "points": P},
        # the potential kernel dominates and is bound by FP64 vector issue,
        # neither HBM nor MFMA
        "roofline": None,
        "kernel_ms_per_step": {k: round(v[0] / a.steps, 4) for k, v in kern.items()},
        "Upot_hartree": U0,
    }
    if not a.no_cpu:
        from oracle import oracle
        done, tc, chunk = 0, 0.0, 512
        while tc < a.cpu_seconds and done < P:
            t0 = time.perf_counter()
            oracle.scp1_average(pr["pes"], q, d, iU, skip + done, chunk)
            tc += time.perf_counter() - t0
            done += chunk
        out["cpu_baseline"] = {"value": done / tc, "unit": "points/s", "cores": 1, "kind": "port",
                               "sample": "%d of the same Sobol points in %d-point chunks (%.1f s): the oracle's "
                                         "serial restatement of SCP1::average (one MPI rank's loop)"
                                         % (done, chunk, tc)}
    print(json.dumps(out))


CONV_DIR = os.path.join(ROOT, "scp3_amd", "data")


def convergence_path(pr, n):
    return os.path.join(CONV_DIR, "convergence_%s_%dstates.json" % (pr["config"], n))


def write_convergence(pr, files, n, merge=False):
    """Convergence record of a workload from --converge outputs: the
    reference levels (mean over scramble seeds of the last checkpoint of the
    stein_mixed runs, with the standard error of that mean) and, for each
    -G form, every checkpoint's largest level deviation from them and the
    first checkpoint from which all levels stay within 0.5 cm^-1 (over
    every seed's checkpoints at and beyond it).  merge: the committed
    record's checkpoints (levels = reference + deviation) and, when the
    outputs hold no stein_mixed run, its reference are kept and the new
    checkpoints added."""
    rows, seen = [], set()

    def add(r):
        key = (r.get("form"), r.get("seed"), r.get("points"), r.get("points_t"))
        # a run continued from an earlier one repeats that run's last checkpoint
        if "points" in r and key not in seen:
            seen.add(key)
            rows.append(r)

    for fn in files:
        for line in open(fn):
            if line.strip().startswith("{"):
                add(json.loads(line))
    old = None
    if merge and os.path.exists(convergence_path(pr, n)):
        old = json.load(open(convergence_path(pr, n)))
        oref = np.array(old["reference"]["levels_cm1"])
        for form, fr in old["forms"].items():
            for t in fr["checkpoints"]:
                add({"form": form, "seed": t["seed"], "points": t["points"], "points_t": t["points_t"],
                     "accumulate_s": t["accumulate_s"], "train_s": t.get("train_s", 0.0),
                     "levels_cm1": [float(v) for v in oref + np.array(t["dev_cm1"])]})
    st = [r for r in rows if r["form"] == "stein_mixed"]
    if st:
        pmax = max(r["points"] for r in st)
        L = np.array([r["levels_cm1"] for r in st if r["points"] == pmax])
        ref = L.mean(0)
        sem = L.std(0, ddof=1) / np.sqrt(L.shape[0]) if L.shape[0] > 1 else np.zeros_like(ref)
        reference = {"estimator": "gradient (Stein), mixed form", "points": pmax,
                     "seeds": sorted(r["seed"] for r in st if r["points"] == pmax),
                     "levels_cm1": [round(float(v), 4) for v in ref],
                     "sem_cm1": [round(float(v), 4) for v in sem]}
    else:
        if old is None:
            raise SystemExit("--converge-summary: no stein_mixed runs for the reference levels and no record to merge")
        reference = old["reference"]
        ref = np.array(reference["levels_cm1"])
    rec = {"workload": "%s, %d states, scrambled Sobol points (skip 2^30)" % (pr["config"], n),
           "reference": reference,
           "forms": {}}
    # the control variate's checkpoints: those of the preset's correction-term size only
    cv_pt = (1 << pr["cv_t"]) if pr.get("cv_t") is not None else None
    rows = [r for r in rows if r["form"] != "cv" or cv_pt is None or r.get("points_t") == cv_pt]
    for form in sorted({r["form"] for r in rows} - {"stein_mixed"}):
        cps = sorted((r for r in rows if r["form"] == form), key=lambda r: (r["points"], r["seed"]))
        table = []
        for r in cps:
            dev = float(np.abs(np.array(r["levels_cm1"]) - ref).max())
            table.append({"points": r["points"], "points_t": r.get("points_t"), "seed": r["seed"],
                          "accumulate_s": r["accumulate_s"], "train_s": r.get("train_s", 0.0),
                          "max_dev_cm1": round(dev, 3),
                          "dev_cm1": [round(float(v), 3) for v in np.array(r["levels_cm1"]) - ref]})
        # the smallest point count from which every checkpoint, of every
        # seed, has all levels within 0.5 cm^-1
        req = None
        for P in sorted({t["points"] for t in table}):
            if all(t["max_dev_cm1"] <= 0.5 for t in table if t["points"] >= P):
                req = P
                break
        rec["forms"][form] = {"checkpoints": table, "points_required": req,
                              "points_t": table[0]["points_t"] if table else None,
                              "dev_at_required_cm1": {str(t["seed"]): t["dev_cm1"] for t in table
                                                      if t["points"] == req} if req else None}
    path = convergence_path(pr, n)
    with open(path, "w") as f:
        json.dump(rec, f, indent=1)
    print("wrote", path, file=sys.stderr)


def time_to_half_wavenumber(pr, res, n):
    """Time to the lowest levels within 0.5 cm^-1 of the converged levels:
    the points each form needs (committed convergence record) at this run's
    measured rates.  The FP64 form is the mixed form's estimator on the same
    points (levels equal to 0.011 cm^-1), so it needs the mixed form's
    points.  None when no record exists for the workload."""
    path = convergence_path(pr, n)
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    out = {"record": os.path.relpath(path, ROOT), "reference_points": rec["reference"]["points"],
           "reference_sem_cm1_max": max(rec["reference"]["sem_cm1"])}
    plain = rec["forms"].get("mixed", {}).get("points_required")
    for f in ("fp64", "mixed"):
        if f in res and plain:
            rr = res[f]
            rate = rr["points"] / rr["t_acc"]
            out[f] = {"points": plain, "points_per_s": rate, "s": plain / rate + rr["t_lev"],
                      "level_error_cm1": rec["forms"]["mixed"]["dev_at_required_cm1"]}
    cvr = rec["forms"].get("cv", {}).get("points_required")
    if "cv" in res and cvr and "mixed" in res:
        rr = res["cv"]
        # this run's step: 2^cv_nn surrogate points plus a correction term
        # over 2^cv_t true points, each a mixed-form point and, when the
        # surrogate takes one product, a surrogate point (the correction
        # subtracts w_nn with that same product)
        rate_t = res["mixed"]["points"] / res["mixed"]["t_acc"]
        both = pr.get("cv_terms", 1) == 1
        Pt_run = 1 << pr["cv_t"]
        rate_nn = ((1 << pr["cv_nn"]) + (Pt_run if both else 0)) / max(rr["t_acc"] - Pt_run / rate_t, 1e-9)
        # the record's correction term may have had more true points
        Pt = rec["forms"]["cv"].get("points_t") or Pt_run
        out["cv"] = {"points": cvr, "points_t": Pt, "surrogate_points_per_s": rate_nn,
                     "train_s": rr["train_s"],
                     "s": rr["train_s"] + (cvr + (Pt if both else 0)) / rate_nn + Pt / rate_t + rr["t_lev"],
                     "level_error_cm1": rec["forms"]["cv"]["dev_at_required_cm1"]}
    if "fp64" in out and "cv" in out:
        out["speedup_cv_vs_fp64"] = out["fp64"]["s"] / out["cv"]["s"]
    if "fp64" in out and "mixed" in out:
        out["speedup_mixed_vs_fp64"] = out["fp64"]["s"] / out["mixed"]["s"]
    return out


def main():
    a = parse()
    if a.copy_bytes:
        copy_calibration(a.copy_bytes)
        return
    if a.gpus != 1:
        print("bench.py: one GPU only; --gpus %d is not supported (the points would be split "
              "across GPUs and the partial H summed once at the end)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import scp3_amd
    from scp3_amd import system

    pr = dict(PRESETS[a.preset])
    if a.points is not None:
        pr["points"] = a.points
    if pr.get("kind") == "scp1":
        bench_scp1(a, pr)
        return
    if a.forms:
        pr["forms"] = a.forms
    if a.tile is not None:
        pr["tile"] = a.tile
    pr["mixed_kernel"] = a.mixed_kernel
    pr["scramble"] = None if a.no_scramble else a.scramble
    sysd = system.prepare(system.input_path(pr["config"]), nmodes2=pr["n2"], nmodes3=pr["n3"])
    sysd["name"] = pr["config"]
    if a.cv_t is not None:
        pr["cv_t"] = a.cv_t
    if a.cv_nn is not None:
        pr["cv_nn"] = a.cv_nn
    if a.mlp_steps is not None:
        pr["mlp_steps"] = a.mlp_steps
    if a.mlp_pairs is not None:
        pr["mlp_pairs"] = a.mlp_pairs
    pr["cv_terms"] = a.cv_terms
    if a.converge_summary:
        # states: ground, singles of the selected modes, doubles of the first
        # n2 and triples of the first n3 (ScaledMatrixElements.cpp:81-99)
        m2, m3 = sysd["nmodes2"], sysd["nmodes3"]
        n = 1 + sysd["modes"].size + m2 * (m2 + 1) // 2 + m3 * (m3 + 1) * (m3 + 2) // 6
        write_convergence(pr, a.converge_summary, n, merge=a.merge_record)
        return
    if a.converge:
        converge(a, pr, sysd)
        return
    P = 1 << pr["points"]
    skip = 1 << 30
    forms = pr["forms"].split(",")
    caps = pr.get("caps", {})
    res, extra, nsteps = {}, {}, {}
    with scp3_amd.Context(0) as c:
        c.L.scp3g_set_cv_terms(c.handle, a.cv_terms)
        for f in forms:
            st, wu = a.steps, a.warmup
            if f in caps and f != forms[0]:
                st, wu = min(st, caps[f][0]), min(wu, caps[f][1])
            nsteps[f] = (st, wu)
            print("bench: preset %s form %s (%d steps, %d warm-up)" % (a.preset, f, st, wu),
                  file=sys.stderr, flush=True)
            res[f] = run_form(c, sysd, pr, f, P, skip, st, wu, pr["levels"], extra,
                              keep_h=f in ("fp64", "mixed") and "fp64" in forms and "mixed" in forms)
            print("bench: form %s accumulate %.4f s, levels %.4f s" % (f, res[f]["t_acc"], res[f]["t_lev"]),
                  file=sys.stderr, flush=True)
    n = res[forms[0]]["n"]
    main_form = "mixed" if "mixed" in forms else ("cv" if "cv" in forms else forms[-1])
    r = res[main_form]
    # dominant kernel: the SYRK's MFMA kernel (split-K partial fold timed
    # apart); algorithmic flops of a symmetric rank-K update: n(n+1) per point
    ms, cnt = r["kernels"]["syrk_mma"]
    avg_s = ms / 1e3 / max(cnt, 1)
    # points_done counts one step; the gradient forms run two SYRKs per tile
    # (FP64) or one over twice the points (mixed)
    pts_per_launch = r["points"] * nsteps[main_form][0] / max(cnt, 1) * (2 if main_form in GRAD_FORMS else 1)
    flops = float(n) * (n + 1) * pts_per_launch
    fp64_main = main_form in ("fp64", "grad", "stein")
    peak = PEAK_FP64_TFLOPS if fp64_main else PEAK_BF16_TFLOPS
    achieved = flops / avg_s / 1e12 if avg_s > 0 else 0.0
    scr = "scrambled (seed %d)" % pr["scramble"] if pr["scramble"] is not None else "unscrambled"
    out = {
        "metric": "Sobol pts/s into H + time to lowest-10 levels at 0.5 cm^-1, GPU forms vs plain",
        "value": r["points"] / r["t_acc"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * (r["t_acc"] + r["t_lev"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp64_main else "f64 H; bf16x3 MFMA SYRK, f32 potential/basis",
        "data": "synthetic: seed-built relaxed cluster (scp3_amd/data), %s Sobol points" % scr,
        "config": {"workload": "%s: %s, %d states, 2^%d Sobol points per step, %d-point tiles, lowest %d levels"
                               % (a.preset, pr["config"], n, pr["points"], round(pts_per_launch), pr["levels"]),
                   "basis_states": n, "points": P, "form": main_form},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": pmc_traffic("k_syrk_x3<3>", pr["config"], n, pts_per_launch)
                     if main_form == "mixed" else None,
                     "kernel": "%s (%s form)" % ("k_syrk_f64_lds" if fp64_main else "k_syrk_x3<3>", main_form),
                     "avg_launch_ms": avg_s * 1e3,
                     "flops_per_launch": flops},
        "forms": {},
    }
    for f, rr in res.items():
        out["forms"][f] = {
            "steps": nsteps[f][0],
            "warmup": nsteps[f][1],
            "points": rr["points"],
            "points_per_s": rr["points"] / rr["t_acc"],
            "time_to_levels_s": rr["t_acc"] + rr["t_lev"] + rr["train_s"],
            "accumulate_s": rr["t_acc"],
            "eigensolve_s": rr["t_lev"],
            "levels_cm1": [round(float(v) * scp3_amd.AUTOCM, 4) for v in rr["levels"]],
            "kernel_ms_per_step": {kk: round(v[0] / nsteps[f][0], 4) for kk, v in rr["kernels"].items() if v[1]},
        }
        if f == "cv":
            out["forms"][f].update(extra)
            out["forms"][f]["points_surrogate"] = 1 << pr["cv_nn"]
            out["forms"][f]["points_t"] = 1 << pr["cv_t"]
    base = main_form.replace("_mixed", "")
    if main_form.endswith("_mixed") and base in res:
        d = np.abs(res[main_form]["levels"] - res[base]["levels"]).max() * scp3_amd.AUTOCM
        out["max_level_diff_vs_%s_cm1" % base] = float(d)
    if "fp64" in res and main_form != "fp64":
        d = np.abs(res[main_form]["levels"] - res["fp64"]["levels"]).max() * scp3_amd.AUTOCM
        out["max_level_diff_vs_fp64_cm1"] = float(d)
    if "fp64" in res and "mixed" in res and res["mixed"]["Mout"] is not None:
        # the mixed form against the FP64 form on the same points
        out["error_budget_mixed_vs_fp64"] = error_budget(res["mixed"]["Mout"], res["fp64"]["Mout"],
                                                         res["mixed"]["levels"], res["fp64"]["levels"])
    for rr in res.values():
        rr.pop("Mout", None)
    tt = time_to_half_wavenumber(pr, res, n)
    if tt:
        out["time_to_0p5_cm1"] = tt
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sysd, "ttm3f" if "ttm3f" in pr["config"] else "qtip4pf", skip,
                                           pr["scramble"], a.cpu_chunk, a.cpu_seconds, P)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
