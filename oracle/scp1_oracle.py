"""SCP1's self-consistent phonon iteration (SCP1.cpp:89-190) and the Hessian
helpers it calls (Hessian.cpp), restated with numpy.

TEST INFRASTRUCTURE ONLY: the oracle that the SCP1 driver's output is
checked against.  The Gaussian averages come from a caller-supplied
function, normally `oracle.scp1_average` (scp1_oracle.c).

Two deliberate differences from the reference, made the same way in the
driver so that both draw the same points:
  * eigenvectors get a canonical sign, their largest-magnitude component
    positive (LAPACK leaves the sign open, and the sign of a column of
    isqrtM_U decides which way its Sobol coordinate displaces the cluster);
  * getHessian (Hessian.cpp:13-37) divides a gradient in kcal/mol/A by a step
    in bohr; here the gradient is converted to Hartree/bohr first.
Pinned by the harmonic closed forms in tests/test_scp1.py.
"""
import numpy as np

BOHR = 0.52917721092          # Constants.cpp:12
AUTOCM = 2.194746313e5        # Constants.cpp:13
AUTOKCALPMOL = 627.509474277194  # Constants.h:21


def disp_disp_corr_diag(omega, kT, cutoff):
    """<q_k^2> of a harmonic mode at temperature kT, zero below the cutoff
    (SCP1.cpp:18-31)."""
    omega = np.asarray(omega, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        if kT == 0.0:
            d = 0.5 / omega
        else:
            d = 0.5 / (omega * np.tanh(0.5 / kT * omega))
    d = np.where(omega < cutoff, 0.0, d)
    return d


def transrot_basis(q, mass):
    """Mass-weighted translations and rotations, orthonormalised
    (Hessian.cpp:89-120)."""
    n = q.size
    U = np.zeros((n, 6))
    qs = q * np.sqrt(mass)
    total = mass.sum() / 3.0
    for i in range(0, n, 3):
        U[i, 0] = np.sqrt(mass[i] / total)
        U[i + 1, 1] = np.sqrt(mass[i + 1] / total)
        U[i + 2, 2] = np.sqrt(mass[i + 2] / total)
        U[i + 1, 3] = -qs[i + 2]
        U[i + 2, 3] = qs[i + 1]
        U[i, 4] = qs[i + 2]
        U[i + 2, 4] = -qs[i]
        U[i, 5] = -qs[i + 1]
        U[i + 1, 5] = qs[i]
    for i in range(3, 6):
        for j in range(i):
            U[:, i] -= U[:, j] * np.dot(U[:, i], U[:, j])
        U[:, i] /= np.linalg.norm(U[:, i])
    return U


def regtransrot(U, H):
    """(1 - U U^T) H (1 - U U^T) (Hessian.cpp:123-130)."""
    HP = H - (H @ U) @ U.T
    return HP - U @ (U.T @ HP)


def mass_scale(mass, H):
    """H(i, j) / sqrt(m_i m_j) (Hessian.cpp:45-55)."""
    s = 1.0 / np.sqrt(mass)
    return H * (s[:, None] * s[None, :])


def eig_sym(K):
    """Ascending eigenpairs of the upper triangle (Armadillo's eig_sym reads
    it), columns signed so that their largest component is positive."""
    w, U = np.linalg.eigh(K, UPLO="U")
    idx = np.argmax(np.abs(U), axis=0)
    sgn = np.sign(U[idx, np.arange(U.shape[1])])
    sgn[sgn == 0] = 1.0
    return w, U * sgn[None, :]


def fd_hessian(q, grad, s=1e-3):
    """Cartesian Hessian (Hartree/bohr^2) by central differences of the
    gradient, step s bohr (Hessian.cpp:13-37); grad(r_angstrom) returns
    (E kcal/mol, dE/dr kcal/mol/A).  Only the upper triangle is
    symmetrised, as in the reference (:38-41)."""
    n = q.size
    r = q * BOHR
    H = np.zeros((n, n))
    for i in range(n):
        r0 = r[i]
        r[i] = r0 + s * BOHR
        gp = np.array(grad(r.copy())[1])
        r[i] = r0 - s * BOHR
        gm = np.array(grad(r.copy())[1])
        H[:, i] = (gp - gm) * BOHR / (2.0 * s * AUTOKCALPMOL)
        r[i] = r0
    for i in range(1, n):
        H[:i, i] = 0.5 * (H[:i, i] + H[i, :i])
    return H


def scp1(q, mass, Ks, average, kT=0.0, max_iterations=200, cutoff=50.0 / AUTOCM, qeps=2.0):
    """SCP1::operator() (SCP1.cpp:89-190) from q (bohr) and the mass-weighted,
    projected Hessian Ks.  average(q, d, isqrtM_U) -> (Upot, UX, UX_q).
    Returns the final q, Ks, omega, F and each iteration's omega, F."""
    q = np.array(q, dtype=np.float64)
    mass = np.asarray(mass, dtype=np.float64)
    Ks = np.array(Ks, dtype=np.float64)
    sm = np.sqrt(mass)
    w, U = eig_sym(Ks)                                   # :105
    omega = np.sqrt(np.abs(w))
    d = disp_disp_corr_diag(omega, kT, cutoff)
    iU = U / sm[:, None]                                 # :109-110
    hist = []
    F = 0.0
    for _ in range(max_iterations):
        Upot, UX, UXq = average(q, d, iU)                # :118
        UXq = UXq * np.sqrt(mass[None, :] / mass[:, None])   # :120-124
        KD = UXq @ U                                     # :126
        for i in range(KD.shape[1]):                     # :127-135
            KD[:, i] = KD[:, i] / d[i] if d[i] != 0.0 else 0.0
        Ks = KD @ U.T                                    # :136
        Ks = 0.5 * (Ks + Ks.T)                           # :138-145
        Ks = regtransrot(transrot_basis(q, mass), Ks)    # :147
        w, U = eig_sym(Ks)                               # :149
        iU = U / sm[:, None]
        omega = np.sqrt(np.abs(w))
        d = disp_disp_corr_diag(omega, kT, cutoff)       # :158
        q = q - qeps * iU @ (d * (iU.T @ UX))            # :160
        F = Upot                                         # :165-173
        oe = omega[6:]
        if kT > 0.0:
            F += kT * np.sum(np.log(2.0 * np.sinh(0.5 / kT * oe))) - 0.25 * np.sum(oe / np.tanh(0.5 / kT * oe))
        else:
            F += 0.25 * np.sum(oe)
        hist.append({"omega": omega.copy(), "F": F, "Upot": Upot, "UX": UX.copy()})
    return {"q": q, "Ks": Ks, "omega": omega, "F": F, "history": hist}
