"""SCP1 on the GPU against the oracle: the Gaussian averages
(scp3g_scp1_average), the finite-difference Hessian (scp3g_hessian) and the
SCP1 driver's iteration (scp3_amd/bin/SCP1 against oracle/scp1_oracle.py)."""
import math
import os
import subprocess

import numpy as np
import pytest

import scp3_amd
from oracle import oracle as O
from oracle import scp1_oracle as S
from scp3_amd import system

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "scp3_amd", "bin")


def load(name):
    species, x, H = system.read_vladimir(system.input_path(name))
    mass, xs, Hs, _ = system.ohhohh(species, x, H)
    return xs / S.BOHR, mass, Hs


def modes(q0, mass, Hs):
    w, U = S.eig_sym(S.regtransrot(S.transrot_basis(q0, mass), Hs))
    d = S.disp_disp_corr_diag(np.sqrt(np.abs(w)), 0.0, 50.0 / S.AUTOCM)
    return d, U / np.sqrt(mass)[:, None]


@pytest.fixture(scope="module")
def ctx():
    c = scp3_amd.Context()
    yield c
    c.close()


@pytest.mark.parametrize("name,pes,npts", [
    ("dimer_qtip4pf", "qtip4pf", (1 << 17) + 1000),   # two tiles, ragged end
    ("hexamer_qtip4pf", "qtip4pf", 5000),
    ("hexamer_ttm3f", "ttm3f", 1500),
    ("w20_ttm3f", "ttm3f", 300),                      # 180 dims: the map in 30-row chunks
])
def test_average_matches_oracle(ctx, name, pes, npts):
    print("scp1 average:", name, pes, npts, flush=True)
    q0, mass, Hs = load(name)
    # off the minimum so that <UX> is not a cancellation
    q = q0 + np.random.default_rng(11).normal(size=q0.size) * 0.01
    d, iU = modes(q0, mass, Hs)
    skip = 1 << (int(math.floor(math.log2(npts))) + 1)
    Ug, UXg, UXqg = ctx.scp1_average(pes, q, d, iU, skip, npts)
    Uo, UXo, UXqo = O.scp1_average(pes, q, d, iU, skip, npts)
    assert abs(Ug - Uo) <= 1e-12 * abs(Uo)
    assert np.abs(UXg - UXo).max() <= 1e-10 * np.abs(UXo).max()
    assert np.abs(UXqg - UXqo).max() <= 1e-10 * np.abs(UXqo).max()


def test_average_rejects_bad_input(ctx):
    q0, mass, Hs = load("dimer_qtip4pf")
    d, iU = modes(q0, mass, Hs)
    bad = d.copy()
    bad[7] = -1.0
    with pytest.raises(scp3_amd.Scp3Error, match="d must be"):
        ctx.scp1_average("qtip4pf", q0, bad, iU, 0, 256)
    with pytest.raises(scp3_amd.Scp3Error, match="point range"):
        ctx.scp1_average("qtip4pf", q0, d, iU, 0, 0)


@pytest.mark.parametrize("name,pes", [("dimer_qtip4pf", "qtip4pf"), ("hexamer_ttm3f", "ttm3f")])
def test_hessian_matches_oracle(ctx, name, pes):
    print("scp1 hessian:", name, pes, flush=True)
    q0, mass, Hs = load(name)
    Hg = ctx.hessian(pes, q0, 1e-3)
    grad = O.qtip4pf_grad if pes == "qtip4pf" else O.ttm3f_grad
    Ho = S.fd_hessian(q0, grad, 1e-3)
    assert np.abs(Hg - Ho).max() <= 1e-8 * np.abs(Ho).max()


def run_scp1(args, cwd):
    return subprocess.run([os.path.join(BIN, "SCP1")] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=300)


def read_coord(path, n3):
    species, x, H = system.read_vladimir(path)
    return x / S.BOHR, H


def check_against_oracle(tmp_path, q0, mass, K0, pes, N, L, tol=1.0, cutoff=50.0, kT=0.0):
    """tol scales every bound: the TTM3-F gradient kernel agrees with the
    oracle to ~1e-9 of the largest component (dipole iterations in another
    order), q-TIP4P/F to rounding."""
    npts = 1 << N
    skip = 1 << (N + 1)
    out = S.scp1(q0, mass, K0, lambda q, d, iU: O.scp1_average(pes, q, d, iU, skip, npts), max_iterations=L,
                 cutoff=cutoff / S.AUTOCM, kT=kT)
    F = np.loadtxt(tmp_path / "free_energy.dat", ndmin=1)
    Fo = np.array([h["F"] for h in out["history"]])
    np.testing.assert_allclose(F, Fo, rtol=1e-10 * tol)
    om = np.loadtxt(tmp_path / "omega_scp1.dat", ndmin=2)
    omo = np.array([h["omega"] for h in out["history"]])
    # modes below the cutoff get d = 0 and leave the effective Hessian with
    # more zero frequencies than the six of translation and rotation: those
    # are rounding noise in both
    live = omo > 1e-6
    np.testing.assert_allclose(om[live], omo[live], rtol=1e-8 * tol)
    assert np.all(om[~live] < 1e-6)
    q, Ks = read_coord(str(tmp_path / "coord.xyz"), q0.size)
    assert np.abs(q - out["q"]).max() < 1e-10 * tol
    assert np.abs(Ks - out["Ks"]).max() < 1e-8 * tol * np.abs(out["Ks"]).max()


def test_scp1_driver_matches_oracle(tmp_path):
    print("scp1 driver: dimer qtip4pf -N 12 -L 3", flush=True)
    path = system.input_path("dimer_qtip4pf")
    r = run_scp1(["-N", "12", "-L", "3", path], str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert "iter: 2" in r.stdout and "F0 = " in r.stdout
    q0, mass, Hs = load("dimer_qtip4pf")
    check_against_oracle(tmp_path, q0, mass, Hs, "qtip4pf", 12, 3)


def test_scp1_driver_without_hessian(tmp_path):
    """No Hessian in the input: the driver starts from the finite-difference
    Hessian (SCP1.cpp:95-102)."""
    print("scp1 driver: dimer qtip4pf, no Hessian, -N 11 -L 2", flush=True)
    species, x, H = system.read_vladimir(system.input_path("dimer_qtip4pf"))
    path = str(tmp_path / "geom.dat")
    with open(path, "w") as f:
        f.write("%d\ngeometry only\n" % len(species))
        for a, s in enumerate(species):
            f.write("%s %.17g %.17g %.17g\n" % (s, x[3 * a], x[3 * a + 1], x[3 * a + 2]))
    r = run_scp1(["-N", "11", "-L", "2", "-S", "5", "-r", "unused.txt", path], str(tmp_path))
    assert r.returncode == 0, r.stderr
    mass, xs, _, _ = system.ohhohh(species, x, None)
    q0 = xs / S.BOHR
    Hc = S.fd_hessian(q0, O.qtip4pf_grad, 1e-3)
    K0 = S.regtransrot(S.transrot_basis(q0, mass), S.mass_scale(mass, Hc))
    check_against_oracle(tmp_path, q0, mass, K0, "qtip4pf", 11, 2)


def test_scp1_driver_ttm3f(tmp_path):
    print("scp1 driver: hexamer ttm3f -N 10 -L 2", flush=True)
    path = system.input_path("hexamer_ttm3f")
    r = run_scp1(["-p", "ttm3f", "-N", "10", "-L", "2", path], str(tmp_path))
    assert r.returncode == 0, r.stderr
    q0, mass, Hs = load("hexamer_ttm3f")
    check_against_oracle(tmp_path, q0, mass, Hs, "ttm3f", 10, 2, tol=100.0)


def test_scp1_reproduces_reference_prism_free_energy(tmp_path):
    """tests/golden/prism_T0K_1M_D50.dat is the reference's own SCP1 result
    for the water hexamer prism (examples/prism_T0K_1M_D50.dat: q-TIP4P/F,
    T = 0 K, 2^20 points, 50 cm^-1 cutoff), whose header holds the free
    energy it reached, 45.5909740976476 kcal/mol.  Twenty iterations from
    its geometry and effective Hessian stay at that fixed point: with
    qeps = 2 the iteration alternates between two states about 0.006
    kcal/mol apart, and the upper one matches the reference's value."""
    print("scp1 driver: reference prism, q-TIP4P/F, -N 20 -L 20", flush=True)
    path = os.path.join(ROOT, "tests", "golden", "prism_T0K_1M_D50.dat")
    with open(path) as f:
        f.readline()
        F_ref = float(f.readline().split("=")[1].split()[0])
    r = run_scp1(["-N", "20", "-L", "20", path], str(tmp_path))
    assert r.returncode == 0, r.stderr
    F = np.loadtxt(tmp_path / "free_energy.dat") * S.AUTOKCALPMOL
    assert F.size == 20
    assert np.all(np.abs(F - F_ref) < 0.01)
    assert abs(F[-10:].max() - F_ref) < 0.002
    species, x, _ = system.read_vladimir(path)
    _, xs, _, _ = system.ohhohh(species, x, None)
    q, _ = read_coord(str(tmp_path / "coord.xyz"), xs.size)
    assert np.abs(q * S.BOHR - xs).max() < 1e-3


def test_scp1_driver_cutoff_and_append(tmp_path):
    """-c sets the frequency cutoff (SCP1_main.cpp:86-88); an input with a
    Hessian opens omega_scp1.dat and free_energy.dat for appending
    (SCP1.h:47,62-65), so a second run adds its iterations."""
    print("scp1 driver: dimer qtip4pf -c 400 -N 10 -L 2, twice", flush=True)
    path = system.input_path("dimer_qtip4pf")
    r = run_scp1(["-c", "400", "-N", "10", "-L", "2", path], str(tmp_path))
    assert r.returncode == 0, r.stderr
    q0, mass, Hs = load("dimer_qtip4pf")
    # modes below 400 cm^-1 leave the effective Hessian with zero eigenvalues
    # beyond the six of translation and rotation; F = <V> + sum(omega[6:])/4
    # (SCP1.cpp:165-173) then adds the square roots of their rounding noise,
    # ~1e-10 Hartree that no two eigensolvers share
    check_against_oracle(tmp_path, q0, mass, Hs, "qtip4pf", 10, 2, tol=300.0, cutoff=400.0)
    r = run_scp1(["-c", "400", "-N", "10", "-L", "1", path], str(tmp_path))
    assert r.returncode == 0, r.stderr
    F = np.loadtxt(tmp_path / "free_energy.dat")
    om = np.loadtxt(tmp_path / "omega_scp1.dat", ndmin=2)
    assert F.size == 3 and om.shape[0] == 3
    # the second run starts again from the input: its first iteration repeats
    assert F[2] == F[0]


def test_scp1_driver_temperature(tmp_path):
    """-T (an addition: the reference's main passes kT = 0) runs
    SCP1::operator() at kT > 0: thermal widths and the harmonic free energy
    (SCP1.cpp:18-31, 165-170).  One iteration: at 300 K the next one
    drops soft modes below the cutoff, and kT log(2 sinh(omega / 2kT)) of
    their zero frequencies is the logarithm of rounding noise, in the
    reference as here."""
    print("scp1 driver: dimer qtip4pf -T 300 -N 11 -L 1", flush=True)
    path = system.input_path("dimer_qtip4pf")
    r = run_scp1(["-T", "300", "-N", "11", "-L", "1", path], str(tmp_path))
    assert r.returncode == 0, r.stderr
    q0, mass, Hs = load("dimer_qtip4pf")
    check_against_oracle(tmp_path, q0, mass, Hs, "qtip4pf", 11, 1, kT=300.0 / 315775.0428639057)
    r = run_scp1(["-T", "-1", path], str(tmp_path))
    assert r.returncode != 0 and "temperature" in r.stderr
