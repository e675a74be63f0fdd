"""TTM3-F: oracle pins (CPU) and the device kernel against the oracle (GPU)."""
import ctypes
import os

import numpy as np
import pytest

import scp3_amd
from oracle import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold():
    return np.load(os.path.join(GOLD, "potentials.npz"), allow_pickle=False)


def _full(crd, tol=1e-15):
    L = oracle.lib()
    L.orc_ttm3f_full.restype = ctypes.c_double
    L.orc_ttm3f_full.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_double,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    c = np.ascontiguousarray(crd, dtype=np.float64)
    nw = c.size // 9
    it = np.zeros(1, dtype=np.int32)
    dip = np.zeros(3 * nw)
    efq = np.zeros(3 * nw)
    T = np.zeros((3 * nw, 3 * nw))
    E = L.orc_ttm3f_full(nw, oracle._p(c), tol, oracle._p(it), oracle._p(dip), oracle._p(efq), oracle._p(T))
    return E, int(it[0]), dip, efq, T


@pytest.mark.parametrize("name", ["dimer_qtip4pf", "hexamer_qtip4pf", "hexamer_ttm3f", "w20_ttm3f"])
def test_oracle_matches_reference(name):
    g = gold()
    if "crd_" + name not in g:
        pytest.skip("no fixture for %s" % name)
    crd = g["crd_" + name]
    for i in range(crd.shape[0]):
        e = oracle.ttm3f(crd[i])
        assert abs(e - g["E_ttm3f_" + name][i]) <= 1e-12 * max(1.0, abs(e)), (i, e)


def test_gammq23_is_incomplete_gamma():
    from scipy.special import gamma, gammaincc
    L = oracle.lib()
    L.orc_gammq23.restype = ctypes.c_double
    L.orc_gammq23.argtypes = [ctypes.c_double]
    for x in np.concatenate([np.linspace(1e-3, 5.0 / 3.0, 40), np.linspace(5.0 / 3.0, 40.0, 80)]):
        assert abs(L.orc_gammq23(x) - gammaincc(2.0 / 3.0, x) * gamma(2.0 / 3.0)) < 2e-13


def test_smear_limits():
    """Smeared kernels approach the bare Coulomb ones at long range."""
    L = oracle.lib()
    L.orc_smear012.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    ts = np.zeros(3)
    r = 12.0
    L.orc_smear012(r, 0.837 ** (1.0 / 3.0), oracle._p(ts))
    np.testing.assert_allclose(ts, [1 / r, 1 / r ** 3, 1 / r ** 5], rtol=1e-12)


@pytest.mark.parametrize("idx", [0, 5])
def test_induced_dipoles_solve_the_linear_system(idx):
    """The mixing iteration converges to mu = polarM (Efq + T mu)."""
    crd = gold()["crd_hexamer_qtip4pf"][idx]
    E, it, dip, efq, T = _full(crd)
    n = dip.size
    direct = np.linalg.solve(np.eye(n) - 1.444 * T, 1.444 * efq)
    np.testing.assert_allclose(dip, direct, rtol=1e-12, atol=1e-12)
    assert 2 < it < 500


def test_cpu_pes_matches_oracle():
    crd = gold()["crd_hexamer_qtip4pf"]
    for i in range(crd.shape[0]):
        assert abs(scp3_amd.pes_cpu("ttm3f", crd[i]) - oracle.ttm3f(crd[i])) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("name,nw", [("dimer_qtip4pf", 2), ("hexamer_qtip4pf", 6), ("hexamer_ttm3f", 6),
                                     ("w20_ttm3f", 20)])
def test_gpu_ttm3f_matches_oracle(gpu_ctx_factory, name, nw):
    torch = pytest.importorskip("torch")
    g = gold()
    if "crd_" + name not in g:
        pytest.skip("no fixture for %s" % name)
    print("ttm3f case", name, flush=True)
    ctx = gpu_ctx_factory(0)
    rng = np.random.default_rng(7)
    base = g["crd_" + name]
    crd = np.concatenate([base, base[:1] + 0.03 * rng.normal(size=(200, base.shape[1]))])
    ref = np.array([oracle.ttm3f(c) for c in crd])
    n = crd.shape[0]
    for fp32 in (0, 1):
        c = torch.from_numpy(np.ascontiguousarray(crd.T)).to(torch.float32 if fp32 else torch.float64).cuda()
        E = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        ctx.check(ctx.L.scp3g_pes_energy(ctx.handle, scp3_amd.PES_TTM3F, nw, n,
                                         ctypes.c_void_p(c.data_ptr()), n, fp32,
                                         ctypes.c_void_p(E.data_ptr())))
        ctx.sync()
        Eg = E.cpu().numpy()
        if fp32:
            err = np.abs(Eg - ref).max()
            bias = (Eg - ref).mean()
            print("ttm3f fp32 max |dE| = %.3e, mean %.2e kcal/mol (%s)" % (err, bias, name))
            # per configuration 5e-3 kcal/mol (1.7 cm^-1); the mean over the
            # configurations, which is what shifts the matrix elements, within
            # 5e-4 kcal/mol (0.17 cm^-1)
            assert err < 5e-3
            assert abs(bias) < 5e-4
        else:
            np.testing.assert_allclose(Eg, ref, rtol=1e-12, atol=1e-11)


# ---- gradients (ttm3f.cpp:325-716) -------------------------------------------

def _ttm3f_geoms(name, count, scale, seed):
    from scp3_amd import system
    sysd = system.prepare(system.input_path(name), nmodes2=0, nmodes3=0)
    rng = np.random.default_rng(seed)
    x0 = sysd["x0"] * scp3_amd.BOHR
    return [x0 + rng.normal(size=x0.size) * scale for _ in range(count)]


@pytest.mark.skipif(not oracle.has_ref(), reason="reference build absent")
@pytest.mark.parametrize("name", ["hexamer_ttm3f", "w20_ttm3f"])
def test_oracle_gradient_matches_reference(name):
    R = oracle.ref()
    for x in _ttm3f_geoms(name, 3, 0.03, 4):
        x = np.ascontiguousarray(x)
        g_ref = np.zeros_like(x)
        e_ref = R.ref_ttm3f_grad(x.size // 9, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(g_ref.ctypes.data))
        e, g = oracle.ttm3f_grad(x)
        assert e == pytest.approx(e_ref, rel=1e-12, abs=1e-10)
        np.testing.assert_allclose(g, g_ref, rtol=0, atol=1e-9 * np.max(np.abs(g_ref)))


def test_oracle_gradient_finite_differences():
    x = _ttm3f_geoms("hexamer_ttm3f", 1, 0.03, 8)[0]
    e, g = oracle.ttm3f_grad(x)
    assert e == pytest.approx(oracle.ttm3f(x), rel=1e-12)
    h = 1e-5
    for i in range(0, x.size, 5):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (oracle.ttm3f(xp) - oracle.ttm3f(xm)) / (2 * h)
        assert abs(fd - g[i]) < 2e-6 * max(1.0, np.max(np.abs(g))), i


def test_cpu_pes_qtip4pf_gradient_is_analytic():
    """The host h2o::PES gradient of q-TIP4P/F equals the oracle's analytic one."""
    from scp3_amd import system
    sysd = system.prepare(system.input_path("hexamer_qtip4pf"), nmodes2=0, nmodes3=0)
    rng = np.random.default_rng(12)
    x = sysd["x0"] * scp3_amd.BOHR + rng.normal(size=sysd["x0"].size) * 0.03
    e, g = scp3_amd.pes_cpu("qtip4pf", x, grad=True)
    e0, g0 = oracle.qtip4pf_grad(x)
    assert e == pytest.approx(e0, rel=1e-12)
    np.testing.assert_allclose(g, g0, rtol=0, atol=1e-11 * np.max(np.abs(g0)))
