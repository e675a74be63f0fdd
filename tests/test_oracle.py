"""Pins of the CPU oracle (oracle/scp3_oracle.c) before it is trusted.

Each part is checked against the reference's own outputs (tests/golden/*.npz,
written by tests/golden/make_golden.py from the reference sources compiled
under oracle/_ref), against oracle/_ref directly when it is built, and
against what the mathematics fixes: the Gray-code form of the Antonov-Saleev
update, (0,k,1)-net stratification, van der Corput on dimension 0, BSM's
published accuracy against scipy's ndtri, the Hermite closed forms, the
basis order of python/base_frequencies.py:42-60, Gauss-Hermite
orthonormality of the product basis, and the harmonic limit of the map.
"""
import itertools
import os

import numpy as np
import pytest

from oracle import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ---- Sobol ----------------------------------------------------------------

def test_sobol_first_2048_all_dims_match_reference():
    g = gold("sobol.npz")["small_dims1111_seed0"]
    np.testing.assert_array_equal(oracle.sobol_ints(1111, 0, 2048), g)


@pytest.mark.parametrize("dims", [12, 48, 174])
def test_sobol_skip30_matches_reference(dims):
    g = gold("sobol.npz")
    np.testing.assert_array_equal(oracle.sobol_ints(dims, 1 << 30, 64), g["ints_skip30_dims%d" % dims])
    # std_normal = i8_sobol + BSM, bit for bit (sobol.cpp:28706-28714)
    np.testing.assert_array_equal(oracle.sobol_normals(dims, 1 << 30, 64, halve=False),
                                  g["normals_skip30_dims%d" % dims])


def test_sobol_unaligned_start_matches_reference():
    g = gold("sobol.npz")["ints_seed123457_dims48"]
    np.testing.assert_array_equal(oracle.sobol_ints(48, 123457, 300), g)


def test_sobol_gray_code_recurrence():
    """Point s equals the XOR of v[j] over the set bits of gray(s) and the
    single-step update v[lo0(s)-1] (brute force on the direction table)."""
    dims = 7
    v = np.zeros((dims, 62), dtype=np.uint64)
    oracle.lib().orc_sobol_table(dims, oracle._p(v))
    pts = oracle.sobol_ints(dims, 0, 1024)
    state = np.zeros(dims, dtype=np.uint64)
    for s in range(1024):
        g = s ^ (s >> 1)
        closed = np.zeros(dims, dtype=np.uint64)
        for j in range(11):
            if (g >> j) & 1:
                closed ^= v[:, j]
        np.testing.assert_array_equal(pts[s], closed)
        np.testing.assert_array_equal(pts[s], state)
        lo0 = (~s & (s + 1)).bit_length()   # 1-based position of the low zero bit
        state = state ^ v[:, lo0 - 1]


def test_sobol_dimension0_is_van_der_corput():
    pts = oracle.sobol_ints(3, 0, 4096)[:, 0]
    for s in range(4096):
        g = s ^ (s >> 1)
        expect = sum(((g >> j) & 1) << (61 - j) for j in range(13))
        assert int(pts[s]) == expect


@pytest.mark.parametrize("dims", [48, 174])
def test_sobol_is_stratified(dims):
    """Every 1-D projection of 2^12 consecutive aligned points hits each of
    the 2^12 dyadic intervals once ((0,12,1)-net property)."""
    for start in (0, 1 << 30):
        top = (oracle.sobol_ints(dims, start, 4096) >> np.uint64(50)).astype(np.int64)
        for d in range(dims):
            assert np.array_equal(np.sort(top[:, d]), np.arange(4096))


# ---- affine digit scrambling -------------------------------------------------

M64 = (1 << 64) - 1


def _splitmix64(st):
    st = (st + 0x9e3779b97f4a7c15) & M64
    z = st
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M64
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M64
    return st, z ^ (z >> 31)


def scramble_ints_py(x, seed):
    """Affine digit scramble L x + e per dimension, bit by bit in Python:
    row i of L (output digit i, digit 0 = bit 61) has bit 61-i set and
    random bits above it from splitmix64 seeded seed ^ (0x9e37... (d+1));
    one more draw gives e."""
    out = np.zeros_like(x)
    allb = (1 << 62) - 1
    for d in range(x.shape[1]):
        st = (seed ^ ((0x9e3779b97f4a7c15 * (d + 1)) & M64)) & M64
        masks = []
        for i in range(62):
            b = 61 - i
            st, r = _splitmix64(st)
            masks.append((1 << b) | (r & allb & ~((2 << b) - 1)))
        st, e = _splitmix64(st)
        e &= allb
        col = []
        for v in x[:, d].tolist():
            o = 0
            for i, m in enumerate(masks):
                o |= (bin(v & m).count("1") & 1) << (61 - i)
            col.append(o ^ e)
        out[:, d] = np.array(col, dtype=np.uint64)
    return out


@pytest.mark.parametrize("seed", [0, 1, 0x5eed])
def test_scramble_restatement_and_net_property(seed):
    """The oracle's affine digit scramble equals the bitwise restatement, is
    affine (s(x) ^ s(y) ^ s(0) = s(x ^ y)), keeps each dimension's first 2^m
    points a permutation of the 2^m dyadic intervals (the unit lower
    triangle maps digit prefixes one to one), and seed 0 is a real scramble."""
    x = oracle.sobol_ints(5, 0, 256)
    sx = oracle.sobol_scramble(x, seed)
    np.testing.assert_array_equal(sx, scramble_ints_py(x, seed))
    rng = np.random.default_rng(seed + 1)
    a = rng.integers(0, 1 << 62, size=(40, 5), dtype=np.uint64)
    b = rng.integers(0, 1 << 62, size=(40, 5), dtype=np.uint64)
    z = oracle.sobol_scramble(np.zeros((1, 5), dtype=np.uint64), seed)
    np.testing.assert_array_equal(oracle.sobol_scramble(a, seed) ^ oracle.sobol_scramble(b, seed) ^ z,
                                  oracle.sobol_scramble(a ^ b, seed))
    for m in (4, 8):
        top = sx[: 1 << m] >> np.uint64(62 - m)
        for d in range(5):
            assert sorted(top[:, d].tolist()) == list(range(1 << m))
    assert np.mean(sx == x) < 0.05
    y = oracle.sobol_normals(5, 0, 256, scramble=seed)
    u = sx.astype(np.float64) * 2.0 ** -62
    np.testing.assert_array_equal(y, oracle.bsm(u.ravel()).reshape(u.shape) / np.sqrt(2.0))


# ---- Beasley-Springer-Moro -------------------------------------------------

def test_bsm_matches_reference_bitwise():
    g = gold("bsm_ho.npz")
    np.testing.assert_array_equal(oracle.bsm(g["bsm_u"]), g["bsm_x"])


def test_bsm_accuracy_and_symmetry():
    from scipy.special import ndtri
    u = np.linspace(1e-9, 1 - 1e-9, 200001)
    x = oracle.bsm(u)
    assert np.abs(x - ndtri(u)).max() < 3.5e-9          # published BSM accuracy
    np.testing.assert_allclose(oracle.bsm(1 - u), -x, atol=2e-15 * (1 + np.abs(x)).max())
    assert oracle.bsm(np.array([0.5]))[0] == 0.0


# ---- basis -------------------------------------------------------------------

def test_hermite_functions():
    from numpy.polynomial.hermite import hermval
    from math import factorial
    g = gold("bsm_ho.npz")
    ys = g["ho_y"]
    L = oracle.lib()
    for n in range(4):
        mine = np.array([L.orc_ho(n, float(v)) for v in ys])
        np.testing.assert_array_equal(mine, g["ho_vals"][n])        # reference bits
        c = np.zeros(n + 1)
        c[n] = 1.0
        closed = hermval(ys, c) / np.sqrt(2.0 ** n * factorial(n))  # normalised H_n
        np.testing.assert_allclose(mine, closed, rtol=1e-14, atol=1e-14)


def state_list(M, M2, M3):
    """Basis order of python/base_frequencies.py:42-60 (index tuples)."""
    sl = [()]
    sl += [(x,) for x in range(M)] + [(x, x) for x in range(M2)]
    for k in range(M2):
        sl += [(k, l) for l in range(k + 1, M2)]
    sl += [(x, x, x) for x in range(M3)]
    for k in range(M3):
        for l in range(k + 1, M3):
            sl += [(k, k, l), (k, l, l)]
    for k in range(M3):
        for l in range(k + 1, M3):
            sl += [(k, l, j) for j in range(l + 1, M3)]
    return sl


@pytest.mark.parametrize("M,M2,M3", [(12, 12, 0), (7, 5, 4), (48, 48, 0), (9, 9, 9)])
def test_get_bra_order_and_values(M, M2, M3):
    from numpy.polynomial.hermite import hermval
    from math import factorial

    def h(n, y):
        c = np.zeros(n + 1)
        c[n] = 1.0
        return hermval(y, c) / np.sqrt(2.0 ** n * factorial(n))

    sl = state_list(M, M2, M3)
    n = oracle.basis_size(M, M2, M3)
    assert n == len(sl)
    rng = np.random.default_rng(5)
    q = rng.normal(scale=0.8, size=(5, M))
    bra = oracle.get_bra(q, M, M2, M3)
    for a, t in enumerate(sl):
        expect = np.ones(5)
        for mode in set(t):
            expect = expect * h(t.count(mode), q[:, mode])
        np.testing.assert_allclose(bra[:, a], expect, rtol=1e-13, atol=1e-14)
    # sub-basis sizes (ScaledMatrixElements.cpp:81-99)
    L = oracle.lib()
    assert L.orc_sub_basis_size(M, M2, M3, 1) == 1 + M
    assert L.orc_sub_basis_size(M, M2, M3, 2) == 1 + M + M2 * (M2 + 1) // 2
    assert L.orc_sub_basis_size(M, M2, M3, 3) == n


@pytest.mark.parametrize("M,M2,M3,sel", [(12, 12, 0, None), (7, 5, 4, None), (9, 9, 9, None),
                                         (12, 4, 2, [0, 1, 2, 3, 7])])
def test_ho_diagonal_brute_force(M, M2, M3, sel):
    """addHODiagonal (ScaledMatrixElements.cpp:735-784) and the zero-point
    energy of unselected modes (SCP3.cpp:238,270), built independently from
    the state list of python/base_frequencies.py:42-60: state t holds
    quanta n_k = t.count(k) and has E0 + sum_k n_k omega_sel[k] with
    E0 = 1/2 sum omega_sel; finalize adds 1/2 (sum omega - sum omega_sel).
    A swapped (k,k,l)/(k,l,l) pair, E0 over all modes or a lost unselected
    ZPE fails one of the cases."""
    rng = np.random.default_rng(M + 10 * M2 + 100 * M3)
    omega = np.sort(rng.uniform(1e-4, 2e-2, size=M))
    modes = np.arange(M, dtype=np.int32) if sel is None else np.array(sel, dtype=np.int32)
    om_sel = omega[modes]
    sl = state_list(modes.size, M2, M3)
    n = len(sl)
    ho = np.array([0.5 * om_sel.sum() + sum(om_sel[k] for k in t) for t in sl])
    D = np.zeros((n, n), order="F")
    oracle.add_ho_diagonal(modes.size, M2, M3, om_sel, D)
    assert not np.any(D - np.diag(np.diag(D)))       # the diagonal only
    np.testing.assert_allclose(np.diag(D), ho, rtol=2e-15, atol=0)
    zpe_rest = 0.5 * (omega.sum() - om_sel.sum())
    if sel is not None:
        assert zpe_rest > 0
    sysd = dict(modes=modes, omega=omega, nmodes2=M2, nmodes3=M3)
    P = 37
    raw = np.asfortranarray(np.triu(rng.normal(size=(n, n))))
    out = oracle.finalize(sysd, raw, P)
    np.testing.assert_allclose(out, raw / P + np.diag(ho + zpe_rest), rtol=1e-14, atol=1e-16)


def test_basis_orthonormal_under_gauss_hermite():
    """<a|b> over |phi_0|^2 ~ exp(-y^2): tensor Gauss-Hermite with 4 nodes per
    mode integrates every product of the degree<=3 basis exactly."""
    M = M2 = M3 = 3
    x, w = np.polynomial.hermite.hermgauss(4)
    w = w / np.sqrt(np.pi)
    pts = np.array(list(itertools.product(x, repeat=M)))
    wts = np.prod(np.array(list(itertools.product(w, repeat=M))), axis=1)
    bra = oracle.get_bra(pts, M, M2, M3)
    n = bra.shape[1]
    S = np.zeros((n, n), order="F")
    oracle.add_epot(bra, wts, S)
    np.testing.assert_allclose(np.triu(S), np.eye(n), atol=1e-13)


def test_add_epot_is_weighted_syrk():
    rng = np.random.default_rng(3)
    bra = rng.normal(size=(300, 40))
    V = rng.normal(size=300)
    M = np.zeros((40, 40), order="F")
    oracle.add_epot(bra, V, M)
    np.testing.assert_allclose(M, np.triu(bra.T @ (V[:, None] * bra)), rtol=1e-12, atol=1e-12)


# ---- potentials ---------------------------------------------------------------

@pytest.mark.parametrize("name", ["dimer_qtip4pf", "hexamer_qtip4pf"])
def test_qtip4pf_matches_reference(name):
    g = gold("potentials.npz")
    crd = g["crd_" + name]
    for i in range(crd.shape[0]):
        e = oracle.qtip4pf(crd[i])
        assert abs(e - g["E_qtip4pf_" + name][i]) <= 1e-12 * max(1.0, abs(e))


def test_qtip4pf_monomer_minimum():
    r, th = 0.9419, np.deg2rad(107.4)
    w = np.array([0, 0, 0, r, 0, 0, r * np.cos(th), r * np.sin(th), 0.0])
    assert abs(oracle.qtip4pf(w)) < 1e-12
    w2 = w.copy()
    w2[3] += 0.01
    assert oracle.qtip4pf(w2) > 0


def test_qtip4pf_gradient_consistency():
    R = oracle.ref()
    if R is None:
        pytest.skip("reference build oracle/_ref not present")
    g = gold("potentials.npz")["crd_hexamer_qtip4pf"][3]
    grad = np.zeros_like(g)
    R.ref_qtip4pf_grad(g.size // 9, oracle._p(g), oracle._p(grad))
    h = 1e-5
    fd = np.array([(oracle.qtip4pf(g + h * e) - oracle.qtip4pf(g - h * e)) / (2 * h)
                   for e in np.eye(g.size)])
    np.testing.assert_allclose(fd, grad, atol=1e-6)


# ---- whole -G block loop --------------------------------------------------------

def test_block_loop_composition(dimer):
    """orc_scp3_accumulate equals its pieces composed in numpy."""
    from scp3_amd import AUTOKCALPMOL, BOHR
    P = 512
    out = oracle.scp3_accumulate(dimer, "qtip4pf", 1 << 30, P, want=("M", "y", "V", "bV"))
    y = oracle.sobol_normals(12, 1 << 30, P)
    np.testing.assert_array_equal(out["y"], y)
    r = BOHR * (dimer["x0"][None, :] + y @ dimer["MUa"].T)
    V = np.array([oracle.qtip4pf(r[i]) for i in range(P)]) / AUTOKCALPMOL
    np.testing.assert_allclose(out["V"], V, rtol=1e-13)
    bV = V - 0.5 * (y * dimer["omega"] * y).sum(1)
    bra = oracle.get_bra(y[:, dimer["modes"]], 12, 12, 0)
    M = np.triu(bra.T @ (bV[:, None] * bra))
    np.testing.assert_allclose(out["M"], M, rtol=1e-11, atol=1e-13)


def test_harmonic_limit_of_the_map(hexamer):
    """V(x0 + MUa y) - V(x0) -> 0.5 sum omega y^2 for small y: pins MUa's
    mass and frequency scaling and the bohr/kcal units of SCP3.cpp:208-223."""
    from scp3_amd import AUTOKCALPMOL, BOHR
    rng = np.random.default_rng(11)
    V0 = oracle.qtip4pf(BOHR * hexamer["x0"])
    for _ in range(5):
        y = 1e-3 * rng.normal(size=48)
        r = BOHR * (hexamer["x0"] + hexamer["MUa"] @ y)
        dV = (oracle.qtip4pf(r) - V0) / AUTOKCALPMOL
        harm = 0.5 * np.sum(hexamer["omega"] * y * y)
        assert abs(dV - harm) < 5e-3 * harm


@pytest.mark.parametrize("name,pes", [("dimer_qtip4pf", "qtip4pf"), ("hexamer_ttm3f", "ttm3f")])
def test_cpu_baseline_blocks(name, pes):
    """The CPU baseline's block pieces (serial potential loop through a C
    function pointer, OpenMP get_bra) equal the oracle's block loop on the
    same scrambled points, with the reference build's potential and with the
    oracle's restatement."""
    from scp3_amd import system
    sysd = system.prepare(system.input_path(name), nmodes2=6, nmodes3=3)
    M0 = sysd["omega"].size
    y = oracle.sobol_normals(M0, 1 << 30, 300, scramble=0)
    o = oracle.scp3_accumulate(sysd, pes, 0, 300, normals=y * np.sqrt(2.0), want=("bV",))
    for reference in (True, False):
        fp, kind = oracle.pes_pointer(pes, reference=reference)
        if reference and kind != "reference":
            continue
        bV = oracle.block_weights(sysd, fp, y)
        np.testing.assert_allclose(bV, o["bV"], rtol=1e-12, atol=1e-15)
    b = oracle.block_bra(sysd, y)
    np.testing.assert_array_equal(b, oracle.get_bra(y[:, sysd["modes"]], M0, 6, 3))
