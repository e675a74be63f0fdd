"""Input preparation for the Hamiltonian-assembly path.

Reads and writes SCP3's "vladimir" input format (DiskIO.cpp:72-116 reader,
DiskIO.cpp:118-140 writer): atom count, a comment line, ``species x y z``
lines in Angstrom, then the mass-weighted Hessian in atomic units as the
upper triangle by columns (H(j, i) for j <= i, one value per line).
Permutes atoms into O H H order (OHHOHH, DiskIO.cpp:142-195) and derives the
normal modes as SCP3's main does (SCP3.cpp:292-318).

The configurations' inputs (water dimer, hexamer, 20-water cluster) are
built here from a seed, relaxed on the FP64 CPU potential and given a
finite-difference Hessian projected free of translations and rotations
(the projection of Hessian.cpp:89-130).  SCP3 itself takes these inputs as
given (it neither relaxes nor computes Hessians).
"""
import os

import numpy as np

from . import AUTOCM, AUTOKCALPMOL, BOHR, HMASS, OMASS, normal_modes, pes_cpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

MASS = {"H": HMASS, "D": 2.0 * HMASS, "O": OMASS, "C": 21891.6543}


def read_vladimir(path):
    """-> species list, x (3N, Angstrom), H (3N x 3N) or None (DiskIO.cpp:72-116)."""
    with open(path) as f:
        lines = f.read().split("\n")
    N = int(lines[0].split()[0])
    species, x = [], []
    for i in range(N):
        t = lines[2 + i].split()
        species.append(t[0])
        x.extend(float(v.replace("D", "E")) for v in t[1:4])
    vals = " ".join(lines[2 + N:]).split()
    n3 = 3 * N
    need = n3 * (n3 + 1) // 2
    H = None
    if len(vals) >= need:
        v = np.array([float(s.replace("D", "E")) for s in vals[:need]])
        H = np.zeros((n3, n3))
        k = 0
        for i in range(n3):
            for j in range(i + 1):
                H[j, i] = H[i, j] = v[k]
                k += 1
    return species, np.array(x), H


def write_vladimir(path, species, x, H, comment="Free energy = 0"):
    """Writer matching read_vladimir (DiskIO.cpp:118-140 layout)."""
    n3 = len(x)
    with open(path, "w") as f:
        f.write("%d\n%s\n" % (len(species), comment))
        for a, s in enumerate(species):
            f.write("%s %.16e %.16e %.16e\n" % (s, x[3 * a], x[3 * a + 1], x[3 * a + 2]))
        for i in range(n3):
            for j in range(i + 1):
                f.write("%.17e\n" % H[j, i])


def write_normals(path, normals):
    """A -r rng file (SCP3.cpp:188-194): one line per point with its nmodes0
    standard normals (not yet divided by sqrt(2)), whitespace-separated."""
    Z = np.atleast_2d(np.asarray(normals, dtype=np.float64))
    with open(path, "w") as f:
        for row in Z:
            f.write(" ".join("%.17g" % v for v in row) + "\n")


def read_normals(path, nmodes0):
    """The normals of a -r rng file, points x nmodes0 (the reference reads
    them in this order with operator>>)."""
    v = np.loadtxt(path, dtype=np.float64).ravel()
    return v[: v.size // nmodes0 * nmodes0].reshape(-1, nmodes0)


def ohhohh(species, x, H=None):
    """Permute coordinates (and H) into O H H blocks (DiskIO.cpp:142-195)."""
    n3 = len(x)
    p = np.zeros(n3, dtype=int)
    iO, iH = 0, 3
    for a, s in enumerate(species):
        i = 3 * a
        if s == "O":
            p[iO:iO + 3] = np.arange(i, i + 3)
            iO += 9
        elif s in ("H", "D"):
            p[iH:iH + 3] = np.arange(i, i + 3)
            iH += 3
            if iH % 9 == 0:
                iH += 3
        else:
            raise ValueError("Uknown mass at i = %d" % i)
    mass = np.repeat([MASS[s] for s in species], 3)[p]
    xs = x[p]
    Hs = H[np.ix_(p, p)] if H is not None else None
    return mass, xs, Hs, p


def prepare(path, nmodes2=None, nmodes3=0, modes=None):
    """Everything SCP3_a receives, from a vladimir file (SCP3.cpp:288-320).

    Returns dict: nw, x0 (bohr, O H H), mass, omega (hartree, all 3N-6 modes),
    MUa (3N x M, column-major), omega_all (cm^-1, omega0.dat), modes, nmodes2,
    nmodes3.  nmodes2 defaults to 3N-6 as SCP3.cpp:305 does.
    """
    species, x, H = read_vladimir(path)
    if H is None:
        raise ValueError("%s holds no Hessian" % path)
    mass, xs, Hs, _ = ohhohh(species, x, H)
    omega, MUa, oall = normal_modes(mass, Hs)
    M = omega.size
    if nmodes2 is None or nmodes2 == 0:
        nmodes2 = M
    if modes is None:
        modes = np.arange(M)
    return dict(nw=len(xs) // 9, x0=xs / BOHR, mass=mass, omega=omega, MUa=MUa,
                omega_all=oall * AUTOCM, modes=np.asarray(modes, dtype=np.int32),
                nmodes2=int(nmodes2), nmodes3=int(nmodes3))


# ---------------------------------------------------------------------------
# cluster construction, relaxation and Hessians (input preparation)


def _water(rng, O):
    """A q-TIP4P/F-like monomer (r 0.9419 A, 107.4 deg) at O, random orientation."""
    r, th = 0.9419, np.deg2rad(107.4)
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    b = rng.normal(size=3)
    b -= a * a.dot(b)
    b /= np.linalg.norm(b)
    h1 = O + r * a
    h2 = O + r * (np.cos(th) * a + np.sin(th) * b)
    return np.concatenate([O, h1, h2])


def build_cluster(nw, seed, d_oo=2.8, d_min=2.6, radius=None):
    """O atoms grown at ~d_oo from a random earlier O, none closer than d_min,
    inside `radius` when given; random H orientations.  O H H order, Angstrom."""
    rng = np.random.default_rng(seed)
    Os = [np.zeros(3)]
    tries = 0
    while len(Os) < nw:
        tries += 1
        if tries > 100000:
            raise RuntimeError("cluster construction failed")
        base = Os[rng.integers(len(Os))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        cand = base + d_oo * v
        if min(np.linalg.norm(cand - o) for o in Os) < d_min - 1e-9:
            continue
        if radius is not None and np.linalg.norm(cand) > radius:
            continue
        Os.append(cand)
    Os = np.array(Os) - np.mean(Os, axis=0)
    return np.concatenate([_water(rng, O) for O in Os])


def relax(x, pes, gtol=1e-7, maxiter=20000):
    from scipy.optimize import minimize

    def f(v):
        E, g = pes_cpu(pes, v, grad=True)
        return E, g

    r = minimize(f, x, jac=True, method="L-BFGS-B",
                 options=dict(gtol=gtol, ftol=1e-15, maxiter=maxiter, maxcor=50))
    return r.x, r.fun, np.abs(r.jac).max()


def hessian(x, pes, h=2e-3):
    """Cartesian Hessian (kcal/mol/A^2) by central differences of the energy."""
    n = len(x)
    H = np.zeros((n, n))
    E0 = pes_cpu(pes, x)
    def E(d):
        return pes_cpu(pes, x + d)
    for i in range(n):
        ei = np.zeros(n)
        ei[i] = h
        H[i, i] = (E(2 * ei) - 2 * E0 + E(-2 * ei)) / (4 * h * h)
        for j in range(i):
            ej = np.zeros(n)
            ej[j] = h
            H[i, j] = H[j, i] = (E(ei + ej) - E(ei - ej) - E(-ei + ej) + E(-ei - ej)) / (4 * h * h)
    return H


def transrot_basis(x, mass):
    """Orthonormal mass-weighted translation/rotation vectors (Hessian.cpp:89-120)."""
    n = len(x)
    q = x * np.sqrt(mass)
    tot = mass.sum() / 3.0
    U = np.zeros((n, 6))
    for i in range(0, n, 3):
        U[i, 0] = np.sqrt(mass[i] / tot)
        U[i + 1, 1] = np.sqrt(mass[i + 1] / tot)
        U[i + 2, 2] = np.sqrt(mass[i + 2] / tot)
        U[i + 1, 3], U[i + 2, 3] = -q[i + 2], q[i + 1]
        U[i, 4], U[i + 2, 4] = q[i + 2], -q[i]
        U[i, 5], U[i + 1, 5] = -q[i + 1], q[i]
    for i in range(3, 6):
        for j in range(i):
            U[:, i] -= U[:, j] * U[:, i].dot(U[:, j])
        U[:, i] /= np.linalg.norm(U[:, i])
    return U


def mass_weighted_hessian(x_ang, H_kcal, mass):
    """kcal/mol/A^2 -> hartree/bohr^2/m_e, mass-weighted, trans/rot projected."""
    Hau = H_kcal * BOHR * BOHR / AUTOKCALPMOL
    s = 1.0 / np.sqrt(mass)
    Hm = Hau * np.outer(s, s)
    U = transrot_basis(x_ang / BOHR, mass)
    HP = Hm - (Hm @ U) @ U.T
    Hp = HP - U @ (U.T @ HP)
    return 0.5 * (Hp + Hp.T)


CONFIGS = {
    # name: (nw, pes, builder kwargs)
    "dimer_qtip4pf": (2, "qtip4pf", dict(d_oo=2.9, d_min=2.6)),
    "hexamer_qtip4pf": (6, "qtip4pf", dict(d_oo=2.8, d_min=2.6)),
    "hexamer_ttm3f": (6, "ttm3f", dict(d_oo=2.8, d_min=2.6)),
    "w20_ttm3f": (20, "ttm3f", dict(d_oo=2.8, d_min=2.6, radius=9.0)),
}


def make_input(name, seed0=0, max_seeds=50, verbose=False):
    """Build, relax and write DATA/<name>.dat; returns (path, seed, frequencies cm^-1)."""
    nw, pes, kw = CONFIGS[name]
    for seed in range(seed0, seed0 + max_seeds):
        x = build_cluster(nw, seed, **kw)
        x, E, gmax = relax(x, pes)
        species = ["O", "H", "H"] * nw
        mass = np.repeat([MASS[s] for s in species], 3)
        Hm = mass_weighted_hessian(x, hessian(x, pes), mass)
        ev = np.linalg.eigvalsh(Hm)
        freq = np.sign(ev) * np.sqrt(np.abs(ev)) * AUTOCM
        ok = freq[6] > 5.0 and np.all(np.abs(freq[:6]) < 1.0)
        if verbose:
            print(name, "seed", seed, "E", E, "gmax", gmax, "lowest", freq[:8])
        if ok:
            os.makedirs(DATA, exist_ok=True)
            path = os.path.join(DATA, name + ".dat")
            write_vladimir(path, species, x, Hm,
                           "E = %.12f kcal/mol (%s, seed %d)" % (E, pes, seed))
            return path, seed, freq
    raise RuntimeError("no minimum found for %s" % name)


def input_path(name):
    return os.path.join(DATA, name + ".dat")


if __name__ == "__main__":
    import sys
    for nm in sys.argv[1:] or ["dimer_qtip4pf", "hexamer_qtip4pf"]:
        print(make_input(nm, verbose=True)[:2])
