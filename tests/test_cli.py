"""The SCP3 and eigensolver drivers (scp3_amd/bin): reference options,
output files, sM_*.h5 round trip."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "scp3_amd", "bin")
DIMER = os.path.join(ROOT, "scp3_amd", "data", "dimer_qtip4pf.dat")


def run(args, cwd):
    return subprocess.run([os.path.join(BIN, args[0])] + args[1:], cwd=cwd, capture_output=True,
                          text=True, timeout=300)


def test_gradient_estimator_unknown_cpu_potential(tmp_path):
    r = run(["SCP3", "-N", "10", "--cpu-pes", "-p", "nonsense", DIMER], str(tmp_path))
    assert r.returncode != 0 and "Unknown H2O potential" in r.stderr


@pytest.mark.parametrize("extra", [[], ["-G", "--cpu-pes"], ["-G", "--cv-true", "99"]])
def test_control_variate_options_rejected(tmp_path, extra):
    """--cv-true needs -G and the device potential, and a sane exponent;
    the driver stops before it touches a GPU."""
    args = ["SCP3", "-N", "10"] + (extra if "99" in extra else extra + ["--cv-true", "10"]) + [DIMER]
    r = run(args, str(tmp_path))
    assert r.returncode != 0 and "--cv-true" in r.stderr


@pytest.mark.gpu
def test_cpu_pes_gradient_run_matches_device_run(tmp_path):
    """--cpu-pes without -G: the gradient estimator from the host potential's
    analytic gradient (q-TIP4P/F) equals the device-potential run."""
    print("cli case: dimer gradient --cpu-pes -N 14", flush=True)
    out = {}
    for tag, extra in (("dev", []), ("cpu", ["--cpu-pes"])):
        d = tmp_path / tag
        d.mkdir()
        r = run(["SCP3", "-N", "14", "-2", "12", "--levels", "5"] + extra + [DIMER], str(d))
        assert r.returncode == 0, r.stderr
        out[tag] = np.loadtxt(d / "levels.dat")
    np.testing.assert_allclose(out["cpu"][:, 0], out["dev"][:, 0], rtol=0, atol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("stein", [False, True])
def test_scp3_gradient_run(tmp_path, stein):
    """SCP3 without -G (the reference's default estimator) against the binding."""
    import scp3_amd
    from scp3_amd import system
    print("cli case: dimer gradient -N 15 stein", stein, flush=True)
    r = run(["SCP3", "-N", "15", "-2", "12", "--levels", "5"] + (["--stein"] if stein else []) + [DIMER],
            str(tmp_path))
    assert r.returncode == 0, r.stderr
    lv = np.loadtxt(tmp_path / "levels.dat")
    e0 = np.loadtxt(tmp_path / "E0_sd.dat")
    sysd = system.prepare(DIMER, nmodes2=12, nmodes3=0)
    with scp3_amd.Context(0) as c:
        c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
        c.set_basis(sysd["modes"], 12, 0)
        c.set_pes("qtip4pf")
        c.set_estimator(scp3_amd.EST_GRADIENT_STEIN if stein else scp3_amd.EST_GRADIENT)
        c.accumulate(1 << 30, 1 << 15)
        ev = c.lowest_levels(5)
        e1 = c.lowest_levels_sub(c.sub_basis_size(1), 1)[0]
    np.testing.assert_allclose(lv[:, 0], ev, rtol=0, atol=1e-12)
    assert abs(e0[-1, 1] - e1 * scp3_amd.AUTOCM) < 1e-6


def test_unknown_potential(tmp_path):
    r = run(["SCP3", "-G", "-p", "nonsense", DIMER], str(tmp_path))
    assert r.returncode != 0 and "Unknown H2O potential" in r.stderr
    r = run(["SCP3", "-G", "-p", "whbb", DIMER], str(tmp_path))
    assert r.returncode != 0 and "WHBB" in r.stderr


def test_missing_input(tmp_path):
    r = run(["SCP3", "-G"], str(tmp_path))
    assert r.returncode != 0 and "No input file" in r.stderr
    r = run(["SCP3", "-G", str(tmp_path / "absent.dat")], str(tmp_path))
    assert r.returncode != 0 and "File not found" in r.stderr


@pytest.mark.gpu
def test_scp3_run_and_eigensolver(tmp_path):
    import scp3_amd
    from scp3_amd import system
    print("cli case: dimer -N 17", flush=True)
    r = run(["SCP3", "-G", "-N", "17", "-2", "12", "--levels", "5", DIMER], str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert len(open(tmp_path / "omega0.dat").read().split()) == 18
    e0 = np.loadtxt(tmp_path / "E0_sd.dat")
    assert e0.shape == (8, 4) and list(e0[:, 0]) == [16384 * (k + 1) for k in range(8)]
    np.testing.assert_allclose(e0[:, 3], e0[:, 2] - e0[:, 1], atol=1e-8)
    assert os.path.exists(tmp_path / "E0_t.dat")
    assert os.path.exists(tmp_path / "sM_0131072.h5")
    lv = np.loadtxt(tmp_path / "levels.dat")
    # the same run through the Python binding
    sysd = system.prepare(DIMER, nmodes2=12, nmodes3=0)
    with scp3_amd.Context(0) as c:
        c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
        c.set_basis(sysd["modes"], 12, 0)
        c.set_pes("qtip4pf")
        c.reset()
        c.accumulate(1 << 30, 1 << 17)
        ev = c.lowest_levels(5)
        e1 = c.lowest_levels_sub(c.sub_basis_size(1), 1)[0]
        Mout = c.h_download()
    np.testing.assert_allclose(lv[:, 0], ev, rtol=0, atol=1e-12)
    assert abs(e0[-1, 1] - e1 * scp3_amd.AUTOCM) < 1e-6
    # the E0 monitor (SCP3.cpp:240-247) against LAPACK on the leading blocks
    # of the same Mout: singles (1 + M) and singles + doubles
    import scipy.linalg
    A = np.triu(Mout) + np.triu(Mout, 1).T
    for col, nsub in ((1, 1 + 12), (2, 1 + 12 + 12 * 13 // 2)):
        lo = scipy.linalg.eigh(A[:nsub, :nsub], eigvals_only=True, subset_by_index=[0, 0])[0]
        assert abs(e0[-1, col] - lo * scp3_amd.AUTOCM) < 1e-6, (col, e0[-1, col], lo * scp3_amd.AUTOCM)
    # eigensolver on the FP32 sM file (eigensolver.cpp:83-109 output format)
    r = run(["eigensolver", "sM_0131072.h5", "5"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    txt = open(tmp_path / "eigenpairs_0131072.dat").read().split("\n")
    assert txt[0] == "# Lowest 5 eigenvalues"
    evs = np.array([float(v) for v in txt[1:6]])
    assert txt[6] == "# Corresponding eigenvectors, one per line."
    np.testing.assert_allclose(evs, ev, rtol=0, atol=2e-7)  # FP32 storage of H
    vec = np.array([float(v) for v in txt[7].split()])
    assert vec.size == 91 and abs(np.linalg.norm(vec) - 1) < 1e-6

def test_cpu_pes_rejects_rng_file_and_unknown_names(tmp_path):
    r = run(["SCP3", "-G", "--cpu-pes", "-p", "nonsense", DIMER], str(tmp_path))
    assert r.returncode != 0 and "Unknown H2O potential" in r.stderr
    r = run(["SCP3", "-G", "--cpu-pes", "-r", "normals.txt", DIMER], str(tmp_path))
    assert r.returncode != 0 and "--cpu-pes" in r.stderr


@pytest.mark.gpu
def test_cpu_pes_run_matches_device_run(tmp_path):
    """--cpu-pes (h2o::PES on the host, the path of WHBB) gives the levels of
    the device-potential run on the same Sobol points."""
    print("cli case: dimer --cpu-pes -N 15", flush=True)
    out = {}
    for tag, extra in (("dev", []), ("cpu", ["--cpu-pes"])):
        d = tmp_path / tag
        d.mkdir()
        r = run(["SCP3", "-G", "-N", "15", "-2", "12", "--levels", "5"] + extra + [DIMER], str(d))
        assert r.returncode == 0, r.stderr
        out[tag] = np.loadtxt(d / "levels.dat")
    np.testing.assert_allclose(out["cpu"][:, 0], out["dev"][:, 0], rtol=0, atol=1e-10)


def test_scp1_options(tmp_path):
    r = run(["SCP1"], str(tmp_path))
    assert r.returncode != 0 and "No input file supplied" in r.stderr
    r = run(["SCP1", "-p", "whbb", DIMER], str(tmp_path))
    assert r.returncode != 0 and "no GPU implementation" in r.stderr
    r = run(["SCP1", "-N", "41", DIMER], str(tmp_path))
    assert r.returncode != 0 and "too large" in r.stderr
    r = run(["SCP1", str(tmp_path / "missing.dat")], str(tmp_path))
    assert r.returncode != 0 and "File not found" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("modes", ["0-3,7", "7,0-3"])
def test_scp3_mode_list_skip_block_and_triples(tmp_path, modes):
    """-m (SCP3.cpp:63-89, the list kept in the given order: -2/-3 take its
    first entries), -S (log2 skip), -B not a multiple of 64 (blocks of 1000
    points until a block ends at or past 2^14: 17000 points, normalised by
    the points summed) and -3, against the binding on the same seeds."""
    import scp3_amd
    from scp3_amd import system
    print("cli case: dimer -m %s -S 20 -B 1000 -2 4 -3 2" % modes, flush=True)
    r = run(["SCP3", "-G", "-N", "14", "-S", "20", "-B", "1000", "-m", modes, "-2", "4", "-3", "2",
             "--levels", "5", DIMER], str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert "5 modes" in r.stderr
    lv = np.loadtxt(tmp_path / "levels.dat")
    sel = np.array([7, 0, 1, 2, 3] if modes.startswith("7") else [0, 1, 2, 3, 7], dtype=np.int32)
    sysd = system.prepare(DIMER, nmodes2=4, nmodes3=2)
    with scp3_amd.Context(0) as c:
        c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
        c.set_basis(sel, 4, 2)
        c.set_pes("qtip4pf")
        assert c.basis_size() == 1 + 5 + 10 + 4
        c.reset()
        c.accumulate(1 << 20, 17000)
        ev = c.lowest_levels(5)
    np.testing.assert_allclose(lv[:, 0], ev, rtol=0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_scp3_control_variate_run(tmp_path, precision):
    """SCP3 -G --cv-true: the surrogate trained on the sampler's pairs, the
    correction term over 2^12 true-potential points held, the surrogate term
    over the -N points; levels.dat against scp3g_accumulate_cv through the
    binding with the same network, training and seeds."""
    import scp3_amd
    from scp3_amd import system
    print("cli case: dimer control variate -N 14 --cv-true 12", precision, flush=True)
    r = run(["SCP3", "-G", "-N", "14", "-2", "12", "--levels", "5", "--precision", precision, "--cv-true", "12",
             "--cv-pairs", "12", "--cv-steps", "100", DIMER], str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert "control variate" in r.stderr
    lv = np.loadtxt(tmp_path / "levels.dat")
    e0 = np.atleast_2d(np.loadtxt(tmp_path / "E0_sd.dat"))   # one checkpoint
    sysd = system.prepare(DIMER, nmodes2=12, nmodes3=0)
    with scp3_amd.Context(0) as c:
        c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
        c.set_basis(sysd["modes"], 12, 0)
        c.set_pes("qtip4pf")
        c.set_precision(scp3_amd.MIXED if precision == "mixed" else scp3_amd.FP64)
        c.mlp_init(256, seed=0)
        c.mlp_sample_data(1 << 40, 1 << 12)
        c.mlp_train(100, 4096, 1e-3)
        c.set_cv_terms(1)
        c.reset()
        c.accumulate_cv(1 << 30, 1 << 14, 1 << 30, 1 << 12)
        ev = c.lowest_levels(5)
        e1 = c.lowest_levels_sub(c.sub_basis_size(1), 1)[0]
    # FP64: the same sums up to rounding; mixed: the binding's single sum
    # scales the weights by 1/P (powers of 4 here) before the BF16 split
    tol = 1e-10 if precision == "fp64" else 1e-8
    np.testing.assert_allclose(lv[:, 0], ev, rtol=0, atol=tol)
    assert abs(e0[-1, 1] - e1 * scp3_amd.AUTOCM) < 1e-6 + tol * scp3_amd.AUTOCM
