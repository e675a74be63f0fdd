"""CPU checks of the SCP1 oracle: the Gaussian averages (oracle/scp1_oracle.c)
and the self-consistent iteration (oracle/scp1_oracle.py).

The reference's SCP1 needs MPI and Armadillo and is not built here; the
oracle is pinned instead by
  (1) the harmonic closed forms of the iteration: for V = 1/2 x^T K x the
      first iteration returns the mass-weighted K whatever Hessian it starts
      from, and the next one gives F = V0 + sum(omega)/2 at T = 0 and
      F = V0 + kT sum log(2 sinh(omega / 2kT)) at T > 0;
  (2) the small-width limit of the averages against the reference build's
      q-TIP4P/F energy and gradient and a finite-difference Hessian of that
      gradient: <V> -> V(q), <UX> -> grad V(q), and
      <UX dq^T> -> Hess V(q) <dq dq^T> + grad V(q) <dq>^T for the sampler's
      dq = isqrtM_U (z o sqrt(d)) (SCP1.cpp:60-62) on the reference's
      normals z;
  (3) the finite-difference Hessian (units fixed) against the input file's
      Hessian, which another implementation (scp3_amd.system, Angstrom
      steps) wrote.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import scp1_oracle as S
from scp3_amd import system


def load(name):
    species, x, H = system.read_vladimir(system.input_path(name))
    mass, xs, Hs, _ = system.ohhohh(species, x, H)
    return xs / S.BOHR, mass, Hs


def ref_qtip4pf():
    """(energy, gradient) of the reference build, else of the oracle (itself
    pinned against the reference build in test_grad.py)."""
    R = O.ref() if O.has_ref() else None
    if R is None:
        return O.qtip4pf_grad

    def f(crd):
        c = np.ascontiguousarray(crd, dtype=np.float64)
        g = np.zeros_like(c)
        e = R.ref_qtip4pf_grad(c.size // 9, O._p(c), O._p(g))
        return e, g
    return f


def harmonic_average(q0, Kc):
    """Exact Gaussian averages of V = 1/2 (x - q0)^T Kc (x - q0)."""
    def avg(q, d, iU):
        Sig = (iU * d[None, :]) @ iU.T
        dx = q - q0
        return 0.5 * dx @ Kc @ dx + 0.5 * np.trace(Kc @ Sig), Kc @ dx, Kc @ Sig
    return avg


@pytest.mark.parametrize("kT", [0.0, 2e-3])
def test_iteration_harmonic_closed_form(kT):
    q0, mass, Hs = load("dimer_qtip4pf")
    T = S.transrot_basis(q0, mass)
    K = S.regtransrot(T, Hs)
    sm = np.sqrt(mass)
    Kc = K * sm[:, None] * sm[None, :]
    # a start that neither equals nor commutes with K, same null space
    rng = np.random.default_rng(3)
    A = rng.normal(size=K.shape) * 1e-3
    K0 = S.regtransrot(T, K + A @ A.T)
    out = S.scp1(q0, mass, K0, harmonic_average(q0, Kc), kT=kT, max_iterations=1)
    assert np.abs(out["Ks"] - K).max() < 1e-9 * np.abs(K).max()
    out = S.scp1(q0, mass, K0, harmonic_average(q0, Kc), kT=kT, max_iterations=2)
    w = np.linalg.eigvalsh(K)
    om = np.sqrt(np.abs(w))[6:]
    np.testing.assert_allclose(out["omega"][6:], om, rtol=1e-9)
    if kT == 0.0:
        F = 0.5 * om.sum()
    else:
        F = kT * np.sum(np.log(2.0 * np.sinh(om / (2.0 * kT))))
    assert abs(out["F"] - F) < 1e-11 * abs(F)
    np.testing.assert_allclose(out["q"], q0, rtol=0, atol=1e-14)


def test_iteration_position_step():
    """q -= qeps isqrtM_U (d o (isqrtM_U^T <UX>)) (SCP1.cpp:160): for a
    harmonic well displaced by delta the step is
    -qeps M^-1/2 U (Omega / 2) U^T M^1/2 delta at T = 0."""
    q0, mass, Hs = load("dimer_qtip4pf")
    T = S.transrot_basis(q0, mass)
    K = S.regtransrot(T, Hs)
    sm = np.sqrt(mass)
    Kc = K * sm[:, None] * sm[None, :]
    w, U = np.linalg.eigh(K)
    delta = (U[:, 6:] @ np.linspace(1.0, 2.0, U.shape[0] - 6)) / sm * 1e-7
    out = S.scp1(q0 + delta, mass, K, harmonic_average(q0, Kc), max_iterations=1)
    Om = np.sqrt(np.abs(w))
    Om[:6] = 0.0
    want = q0 + delta - 2.0 * ((U * (Om / 2.0)[None, :]) @ (U.T @ (sm * delta))) / sm
    assert np.abs(out["q"] - want).max() < 1e-4 * np.abs(delta).max()


def test_fd_hessian_matches_input_hessian():
    q0, mass, Hs = load("dimer_qtip4pf")
    Hc = S.fd_hessian(q0, ref_qtip4pf())
    Hc = np.triu(Hc) + np.triu(Hc, 1).T
    K = S.regtransrot(S.transrot_basis(q0, mass), S.mass_scale(mass, Hc))
    K_in = S.regtransrot(S.transrot_basis(q0, mass), Hs)
    # the two step sizes (1e-3 bohr here, 2e-3 A there) leave ~5e-5 of
    # truncation error; a missing bohr factor would leave 47%
    assert np.abs(K - K_in).max() < 2e-4 * np.abs(K_in).max()


def modes(q0, mass, Hs, scale=1.0):
    w, U = S.eig_sym(S.regtransrot(S.transrot_basis(q0, mass), Hs))
    om = np.sqrt(np.abs(w))
    d = S.disp_disp_corr_diag(om, 0.0, 50.0 / S.AUTOCM) * scale
    return d, U / np.sqrt(mass)[:, None]


def test_average_small_width_limit():
    q0, mass, Hs = load("dimer_qtip4pf")
    n3 = q0.size
    skip, npts = 1 << 13, 1 << 12
    f = ref_qtip4pf()
    # away from the minimum, where the gradient is large
    q0 = q0 + np.random.default_rng(5).normal(size=n3) * 0.02
    E0, g0 = f(q0 * S.BOHR)
    g0 = g0 * S.BOHR / S.AUTOKCALPMOL
    # vanishing widths: <V> and <UX> at q
    d, iU = modes(q0, mass, Hs, 1e-14)
    Up, UX, UXq = O.scp1_average("qtip4pf", q0, d, iU, skip, npts)
    assert abs(Up - E0 / S.AUTOKCALPMOL) < 1e-9 * abs(E0 / S.AUTOKCALPMOL)
    assert np.abs(UX - g0).max() < 1e-6 * np.abs(g0).max()
    # small widths: the linear response through the Hessian
    d, iU = modes(q0, mass, Hs, 1e-6)
    Up, UX, UXq = O.scp1_average("qtip4pf", q0, d, iU, skip, npts)
    z = O.sobol_normals(n3, skip, npts, halve=False)
    dq = (z * np.sqrt(d)[None, :]) @ iU.T
    Sig = dq.T @ dq / npts
    mu = dq.mean(axis=0)
    Hc = S.fd_hessian(q0, f)
    Hc = np.triu(Hc) + np.triu(Hc, 1).T
    want = Hc @ Sig + np.outer(g0, mu)
    assert np.abs(UXq - want).max() < 1e-3 * np.abs(Hc @ Sig).max()
    # a transposed moment (dq g^T) is far off: the check has teeth
    assert np.abs(UXq.T - want).max() > 0.1 * np.abs(Hc @ Sig).max()


def test_average_zero_points_rejected_by_shape():
    q0, mass, Hs = load("dimer_qtip4pf")
    d, iU = modes(q0, mass, Hs)
    Up, UX, UXq = O.scp1_average("qtip4pf", q0, d, iU, 1 << 5, 16)
    assert np.isfinite(Up) and UX.shape == (18,) and UXq.shape == (18, 18)
