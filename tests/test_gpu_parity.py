"""HIP path vs the CPU oracle, through the C ABI (libscp3g.so).

Bars: Sobol integers and basis values bit-exact; normals within the device
log's ulps; map and FP64 potential to 1e-12 relative; FP64 H within 1e-9 of
max|H| (and Frobenius) against the oracle on the same points; mixed H within
the stated split-BF16 tolerance; levels within 1e-8 hartree of LAPACK on the
same matrix.  Sizes are small enough for the oracle to finish in seconds.
"""
import numpy as np
import pytest

import scp3_amd
from oracle import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def vp(t):
    """Device pointer for the C ABI.  The library works on its own stream,
    which does not wait for torch's: finish torch's pending work first."""
    import ctypes
    torch.cuda.synchronize()
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def ctx(gpu_ctx_factory):
    return gpu_ctx_factory(0)


def setup_system(c, sysd, pes="qtip4pf", precision=scp3_amd.FP64, tile=4096):
    c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
    c.set_basis(sysd["modes"], sysd["nmodes2"], sysd["nmodes3"])
    c.set_pes(pes)
    c.set_precision(precision)
    c.set_tile(tile)


# ---- Sobol + BSM ---------------------------------------------------------------

@pytest.mark.parametrize("dims,skip,npts", [(48, 1 << 30, 4096), (12, 0, 3000), (174, 123457, 1111),
                                            (1111, 0, 2048)])
def test_sobol_integers_bit_exact(ctx, dims, skip, npts):
    print("sobol case", dims, skip, npts, flush=True)
    ld = (npts + 255) // 256 * 256
    y = torch.zeros(dims * ld, dtype=torch.float64, device="cuda")
    q = torch.zeros(dims * ld, dtype=torch.int64, device="cuda")
    ctx.check(ctx.L.scp3g_sobol_normals(ctx.handle, dims, skip, npts, 0, vp(y), ld, vp(q)))
    ctx.sync()
    got = q.cpu().numpy().view(np.uint64).reshape(dims, ld)[:, :npts].T
    np.testing.assert_array_equal(got, oracle.sobol_ints(dims, skip, npts))
    yn = y.cpu().numpy().reshape(dims, ld)[:, :npts].T
    ref = oracle.sobol_normals(dims, skip, npts, halve=False)
    # central branch is bit-exact; the tail branch uses the device log()
    assert np.mean(yn == ref) > 0.8
    np.testing.assert_allclose(yn, ref, rtol=4e-15, atol=0)


def test_sobol_halved_matches_golden(ctx):
    g = np.load("tests/golden/sobol.npz")["normals_skip30_dims48"]
    y = torch.zeros(48 * 256, dtype=torch.float64, device="cuda")
    ctx.check(ctx.L.scp3g_sobol_normals(ctx.handle, 48, 1 << 30, 64, 1, vp(y), 256, None))
    ctx.sync()
    yn = y.cpu().numpy().reshape(48, 256)[:, :64].T
    np.testing.assert_allclose(yn, g / np.sqrt(2.0), rtol=4e-15, atol=0)


# ---- map + potential --------------------------------------------------------------

@pytest.mark.parametrize("fp32", [0, 1])
def test_map_and_qtip4pf(ctx, hexamer, fp32):
    setup_system(ctx, hexamer)
    P, ld = 1000, 1024
    yh = oracle.sobol_normals(48, 1 << 30, P)
    ydev = np.zeros((48, ld))
    ydev[:, :P] = yh.T
    y = dev(ydev)
    crd = torch.zeros(54 * ld, dtype=torch.float32 if fp32 else torch.float64, device="cuda")
    ctx.check(ctx.L.scp3g_map_coords(ctx.handle, vp(y), ld, P, vp(crd), ld, fp32))
    E = torch.zeros(P, dtype=torch.float64, device="cuda")
    ctx.check(ctx.L.scp3g_pes_energy(ctx.handle, scp3_amd.PES_QTIP4PF, 6, P, vp(crd), ld, fp32, vp(E)))
    ctx.sync()
    r_ref = scp3_amd.BOHR * (hexamer["x0"][None, :] + yh @ hexamer["MUa"].T)
    r = crd.cpu().numpy().astype(np.float64).reshape(54, ld)[:, :P].T
    E_ref = np.array([oracle.qtip4pf(r_ref[i]) for i in range(P)])
    Eg = E.cpu().numpy()
    if fp32:
        np.testing.assert_allclose(r, r_ref, atol=4e-7 * np.abs(r_ref).max())
        # Rounding the coordinates to FP32 alone moves E by up to ~5e-4
        # kcal/mol on this cluster (oracle on rounded vs exact coordinates),
        # without bias.  Bounds: end to end 1e-3 kcal/mol (< 1.4e-3 = 0.5
        # cm^-1) with |mean| < 2e-5; the kernel's own FP32 arithmetic (135
        # Coulomb terms of O(100) kcal/mol, summed in double), on the same
        # FP32 coordinates, 5e-4.
        d = Eg - E_ref
        assert np.abs(d).max() < 1e-3 and abs(d.mean()) < 2e-5, (np.abs(d).max(), d.mean())
        E_own = np.array([oracle.qtip4pf(r[i]) for i in range(P)])
        assert np.abs(Eg - E_own).max() < 5e-4, np.abs(Eg - E_own).max()
    else:
        np.testing.assert_allclose(r, r_ref, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(Eg, E_ref, rtol=1e-12, atol=1e-11)


def test_qtip4pf_golden_configs(ctx):
    g = np.load("tests/golden/potentials.npz")
    for name, nw in (("dimer_qtip4pf", 2), ("hexamer_qtip4pf", 6)):
        crd = g["crd_" + name]
        n = crd.shape[0]
        c = dev(np.ascontiguousarray(crd.T))  # [coord][config]
        E = torch.zeros(n, dtype=torch.float64, device="cuda")
        ctx.check(ctx.L.scp3g_pes_energy(ctx.handle, scp3_amd.PES_QTIP4PF, nw, n, vp(c), n, 0, vp(E)))
        ctx.sync()
        np.testing.assert_allclose(E.cpu().numpy(), g["E_qtip4pf_" + name], rtol=1e-12, atol=1e-11)


# ---- basis table ---------------------------------------------------------------------

@pytest.mark.parametrize("M2,M3,npts", [(48, 0, 1000), (10, 10, 300), (48, 12, 257)])
def test_basis_bit_exact(ctx, hexamer, M2, M3, npts):
    ctx.set_system(hexamer["x0"], hexamer["omega"], hexamer["MUa"])
    modes = np.arange(48, dtype=np.int32)
    ctx.set_basis(modes, M2, M3)
    n = ctx.basis_size()
    ld = (npts + 255) // 256 * 256
    rows = (n + 63) // 64 * 64
    yh = oracle.sobol_normals(48, 1 << 30, npts)
    ydev = np.zeros((48, ld))
    ydev[:, :npts] = yh.T
    y = dev(ydev)
    for fp32 in (0, 1):
        B = torch.full((rows * ld,), 7.0, dtype=torch.float32 if fp32 else torch.float64, device="cuda")
        ctx.check(ctx.L.scp3g_basis_table(ctx.handle, vp(y), ld, npts, vp(B), ld, fp32))
        ctx.sync()
        Bh = B.cpu().numpy().astype(np.float64).reshape(rows, ld)
        ref = oracle.get_bra(yh, 48, M2, M3)
        if fp32:
            np.testing.assert_array_equal(Bh[:n, :npts], ref.T.astype(np.float32).astype(np.float64))
        else:
            np.testing.assert_array_equal(Bh[:n, :npts], ref.T)
        assert np.all(Bh[:n, npts:] == 0) and np.all(Bh[n:, :] == 0)


# ---- SYRK ---------------------------------------------------------------------------------

def _syrk_case(ctx, n, P, mode):
    rng = np.random.default_rng(n + P)
    ld = (P + 255) // 256 * 256
    rows = (n + 127) // 128 * 128
    B = np.zeros((rows, ld))
    B[:n, :P] = rng.normal(size=(n, P))
    w = np.zeros(ld)
    w[:P] = rng.normal(size=P) - 0.3
    H = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    Hc = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    if mode == scp3_amd.MIXED:
        Bd = dev(B.astype(np.float32))
    else:
        Bd = dev(B)
    wd = dev(w)
    ctx.check(ctx.L.scp3g_h_accumulate(ctx.handle, mode, n, vp(Bd), ld, vp(wd), P, vp(H), vp(Hc)))
    ctx.sync()
    Hh = H.cpu().numpy().reshape(n, n).T  # column-major
    Bn = B[:n, :P]
    if mode == scp3_amd.MIXED:
        Bn = Bn.astype(np.float32).astype(np.float64)
    ref = np.triu(Bn @ (w[:P, None] * Bn.T))
    return Hh, ref


@pytest.mark.parametrize("n,P", [(91, 4096), (1225, 4096), (100, 300), (64, 16), (1300, 8192)])
def test_syrk_fp64(ctx, n, P):
    Hh, ref = _syrk_case(ctx, n, P, scp3_amd.FP64)
    assert np.abs(Hh - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(np.tril(Hh, -1) == 0)


@pytest.mark.parametrize("n,P", [(91, 4096), (1225, 4096), (100, 300)])
def test_syrk_mixed(ctx, n, P):
    Hh, ref = _syrk_case(ctx, n, P, scp3_amd.MIXED)
    # 3-term split-BF16 products (relative 2^-16 per term) accumulated in FP32
    err = np.abs(Hh - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err
    assert np.all(np.tril(Hh, -1) == 0)


# ---- whole path ---------------------------------------------------------------------------

def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max(), np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("skip,P", [(1 << 30, 1 << 14), (12345, 5000)])
def test_accumulate_dimer_fp64(ctx, dimer, skip, P):
    print("dimer case", skip, P, flush=True)
    setup_system(ctx, dimer, tile=2048)
    ctx.reset()
    ctx.accumulate(skip, P)
    M = ctx.h_download(raw=True)
    ref = oracle.scp3_accumulate(dimer, "qtip4pf", skip, P)["M"]
    mx, fro = _rel(M, ref)
    assert mx < 1e-9 and fro < 1e-9, (mx, fro)
    Mout = ctx.h_download()
    np.testing.assert_allclose(Mout, oracle.finalize(dimer, ref, P), rtol=0, atol=1e-9 * np.abs(ref).max() / P)


def test_accumulate_hexamer_fp64_and_levels(ctx, hexamer):
    P = 4096 + 1000
    setup_system(ctx, hexamer)
    ctx.reset()
    ctx.accumulate(1 << 30, P)
    M = ctx.h_download(raw=True)
    ref = oracle.scp3_accumulate(hexamer, "qtip4pf", 1 << 30, P)["M"]
    mx, fro = _rel(M, ref)
    assert mx < 1e-9 and fro < 1e-9, (mx, fro)
    # levels on the GPU vs LAPACK on the same Mout
    import scipy.linalg
    Mout = ctx.h_download()
    A = np.triu(Mout) + np.triu(Mout, 1).T
    ev_ref = scipy.linalg.eigh(A, eigvals_only=True, subset_by_index=[0, 9])
    ev = ctx.lowest_levels(10)
    np.testing.assert_allclose(ev, ev_ref, rtol=0, atol=1e-8)


def test_accumulate_normals_matches_sobol_path(ctx, dimer):
    setup_system(ctx, dimer, tile=1024)
    P = 3000
    y = oracle.sobol_normals(12, 1 << 30, P, halve=False)
    ctx.reset()
    ctx.accumulate_normals(y)
    M1 = ctx.h_download(raw=True)
    ref = oracle.scp3_accumulate(dimer, "qtip4pf", 0, P, normals=y)["M"]
    mx, _ = _rel(M1, ref)
    assert mx < 1e-9


@pytest.mark.parametrize("mode", [scp3_amd.FP64, scp3_amd.MIXED])
def test_harmonic_potential_identity(ctx, hexamer, mode):
    """Harmonic-PES identity (SCP3.cpp:219-223,237-239,269-271): energies
    E = autokcalpmol * 1/2 y.(omega o y) make every weight bV zero, so Mout
    is the HO diagonal plus the unselected modes' zero-point energy and the
    levels are E0 + sum_k n_k omega_k.  The diagonal is built here from the
    state list, not from the oracle.  A -m subset of 12 of the 48 modes with
    doubles and triples (147 states) covers every diagonal family."""
    from test_oracle import state_list
    print("harmonic identity precision", mode, flush=True)
    modes = np.array(list(range(10)) + [20, 30], dtype=np.int32)
    ctx.set_system(hexamer["x0"], hexamer["omega"], hexamer["MUa"])
    ctx.set_basis(modes, 12, 6)
    ctx.set_pes("qtip4pf")
    ctx.set_precision(mode)
    ctx.set_tile(2048)
    skip, P = 1 << 30, 5000
    y = oracle.sobol_normals(48, skip, P)
    E = scp3_amd.AUTOKCALPMOL * 0.5 * (y * hexamer["omega"] * y).sum(1)
    ctx.reset()
    ctx.accumulate_energies(skip, E)
    om = hexamer["omega"]
    om_sel = om[modes]
    sl = state_list(modes.size, 12, 6)
    ho = np.array([0.5 * om_sel.sum() + sum(om_sel[k] for k in t) for t in sl])
    ho += 0.5 * (om.sum() - om_sel.sum())
    Mout = ctx.h_download()
    assert Mout.shape == (len(sl), len(sl)) == (147, 147)
    assert np.abs(Mout - np.diag(ho)).max() < 1e-14 * ho.max()
    ev = ctx.lowest_levels(10, 1e-11)
    np.testing.assert_allclose(ev, np.sort(ho)[:10], rtol=0, atol=1e-12)


def test_mixed_matches_fp64_within_budget(ctx, hexamer):
    P = 1 << 15
    out = {}
    for mode in (scp3_amd.FP64, scp3_amd.MIXED):
        setup_system(ctx, hexamer, precision=mode)
        ctx.reset()
        ctx.accumulate(1 << 30, P)
        out[mode] = (ctx.h_download(), ctx.lowest_levels(10))
    (H64, ev64), (Hmx, evmx) = out[scp3_amd.FP64], out[scp3_amd.MIXED]
    dH = np.abs(Hmx - H64).max()
    dlev = np.abs(evmx - ev64).max() * scp3_amd.AUTOCM
    print("mixed vs fp64: max|dH| %.3e Eh, max|dlevel| %.4f cm^-1" % (dH, dlev))
    assert dlev < 0.5


@pytest.mark.parametrize("n2,n3", [(48, 0), (48, 12)])
def test_mixed_kernels_agree(ctx, hexamer, n2, n3):
    """The signed 256-tile split-BF16 SYRK and the 64-tile one give the same
    H to FP32 accumulation accuracy, and both stay close to the FP64 form."""
    P = 3 * 4096 + 500
    out = {}
    for name, mode, fast in (("fp64", scp3_amd.FP64, 1), ("slow", scp3_amd.MIXED, 0),
                             ("fast", scp3_amd.MIXED, 1)):
        ctx.set_system(hexamer["x0"], hexamer["omega"], hexamer["MUa"])
        ctx.set_basis(np.arange(48, dtype=np.int32), n2, n3)
        ctx.set_pes("qtip4pf")
        ctx.set_precision(mode)
        ctx.set_mixed_kernel(fast)
        ctx.set_tile(4096)
        ctx.reset()
        ctx.accumulate(1 << 30, P)
        out[name] = ctx.h_download()
    ctx.set_mixed_kernel(1)
    scale = np.abs(out["fp64"]).max()
    for k in ("fast",):
        d_fs = np.abs(out[k] - out["slow"]).max() / scale
        d_f64 = np.abs(out[k] - out["fp64"]).max() / scale
        print("n=%d %s: |-slow| %.2e  |-fp64| %.2e (relative to max|H|)" % (out["fp64"].shape[0], k, d_fs, d_f64))
        assert d_fs < 2e-5 and d_f64 < 2e-5
        assert np.all(np.tril(out[k], -1) == 0)


def test_max_basis_20825_states(ctx, hexamer):
    """The largest basis of the workloads (total quanta <= 3 over 48 modes,
    20825 states, 3.47 GB of FP64 H) on a ragged point count: FP64 H against
    the oracle on sampled elements (all points, in chunks), the mixed form
    against the FP64 form on the whole upper triangle, lower triangles zero."""
    rng = np.random.default_rng(11)
    P = (1 << 15) + 123
    out = {}
    for mode in (scp3_amd.FP64, scp3_amd.MIXED):
        print("max-basis case precision", mode, flush=True)
        ctx.set_system(hexamer["x0"], hexamer["omega"], hexamer["MUa"])
        ctx.set_basis(hexamer["modes"], 48, 48)
        ctx.set_pes("qtip4pf")
        ctx.set_precision(mode)
        ctx.set_tile(0)
        ctx.reset()
        ctx.accumulate(1 << 30, P)
        out[mode] = ctx.h_download(raw=True)
    H64, Hmx = out[scp3_amd.FP64], out[scp3_amd.MIXED]
    n = H64.shape[0]
    assert n == 20825
    # sampled elements: random pairs, the diagonal and the last rows/columns
    ii = rng.integers(0, n, 3000)
    jj = rng.integers(0, n, 3000)
    ii, jj = np.minimum(ii, jj), np.maximum(ii, jj)
    d = rng.integers(0, n, 300)
    ii = np.concatenate([ii, d, np.arange(n - 40, n), rng.integers(0, n, 40)])
    jj = np.concatenate([jj, d, np.full(40, n - 1), np.arange(n - 40, n)])
    ii, jj = np.minimum(ii, jj), np.maximum(ii, jj)
    ref = np.zeros(ii.size)
    skip, chunk = 1 << 30, 4096
    for p0 in range(0, P, chunk):
        m = min(chunk, P - p0)
        y = oracle.sobol_normals(48, skip + p0, m)
        r = scp3_amd.BOHR * (hexamer["x0"][None, :] + y @ hexamer["MUa"].T)
        V = np.array([oracle.qtip4pf(r[k]) for k in range(m)]) / scp3_amd.AUTOKCALPMOL
        bV = V - 0.5 * (y * hexamer["omega"] * y).sum(1)
        bra = oracle.get_bra(y[:, hexamer["modes"]], 48, 48, 48)
        ref += np.einsum("pk,p,pk->k", bra[:, ii], bV, bra[:, jj])
        del bra
    got = H64[ii, jj]  # upper triangle: row ii <= column jj
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max() / scale
    print("max-basis fp64 vs oracle on %d elements: %.2e" % (ii.size, err))
    assert err < 1e-9
    scale = np.abs(H64).max()
    dmx = np.abs(Hmx - H64).max() / scale
    print("max-basis mixed vs fp64: max|dH|/max|H| %.2e" % dmx)
    assert dmx < 2e-5
    for H in (H64, Hmx):
        assert not np.any(np.tril(H, -1))


def test_w20_ttm3f_mixed_matches_fp64(ctx):
    """20-water cluster under TTM3-F, all 174 modes selected (the mixed basis
    kernel then stages 174 modes of Hermite values, 134 KiB of LDS), doubles
    over the first 20 modes (385 states): the mixed form's H and lowest
    levels against the FP64 form on the same points, and the direct-load
    mixed SYRK against the LDS-tiled one."""
    from scp3_amd import system
    w20 = system.prepare(system.input_path("w20_ttm3f"), nmodes2=20, nmodes3=0)
    P = 3 * 4096 + 77
    out = {}
    for name, mode, fast in (("fp64", scp3_amd.FP64, 1), ("slow", scp3_amd.MIXED, 0), ("fast", scp3_amd.MIXED, 1)):
        print("w20 case", name, flush=True)
        setup_system(ctx, w20, pes="ttm3f", precision=mode)
        ctx.set_mixed_kernel(fast)
        ctx.reset()
        ctx.accumulate(1 << 30, P)
        out[name] = (ctx.h_download(), ctx.lowest_levels(5))
    ctx.set_mixed_kernel(1)
    H64, ev64 = out["fp64"]
    assert H64.shape == (385, 385)
    scale = np.abs(H64).max()
    for k in ("slow", "fast"):
        H, ev = out[k]
        dH = np.abs(H - H64).max() / scale
        dlev = np.abs(ev - ev64).max() * scp3_amd.AUTOCM
        print("w20 %s vs fp64: max|dH|/max|H| %.2e, max|dlevel| %.4f cm^-1" % (k, dH, dlev))
        assert dH < 3e-4 and dlev < 0.5
    assert np.abs(out["fast"][0] - out["slow"][0]).max() / scale < 2e-5


def test_cpu_potential_path_matches_device_path(ctx, dimer):
    """A potential evaluated by the caller (here the oracle's q-TIP4P/F on
    the CPU, standing in for WHBB behind h2o::PES) through
    scp3g_points_coords + scp3g_accumulate_energies gives the H of the
    device-potential path on the same Sobol points."""
    setup_system(ctx, dimer, tile=1024)
    skip, P = 1 << 30, 2500
    crd = ctx.points_coords(skip, P)
    r_ref = scp3_amd.BOHR * (dimer["x0"][None, :] + oracle.sobol_normals(12, skip, P) @ dimer["MUa"].T)
    np.testing.assert_allclose(crd, r_ref, rtol=1e-13, atol=1e-13)
    E = np.array([oracle.qtip4pf(crd[i]) for i in range(P)])
    ctx.reset()
    ctx.accumulate_energies(skip, E)
    H_cpu = ctx.h_download(raw=True)
    ctx.reset()
    ctx.accumulate(skip, P)
    H_dev = ctx.h_download(raw=True)
    mx, fro = _rel(H_cpu, H_dev)
    assert mx < 1e-11 and fro < 1e-11, (mx, fro)


def test_two_contexts_in_two_threads(dimer, hexamer):
    """One context per host thread (the C ABI's threading rule); the tile and
    work-order caches are shared by the contexts of a process."""
    import threading
    out, errs = {}, []

    def work(tag, sysd, mode):
        try:
            with scp3_amd.Context(0) as c:
                c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
                c.set_basis(sysd["modes"], sysd["nmodes2"], sysd["nmodes3"])
                c.set_pes("qtip4pf")
                c.set_precision(mode)
                c.set_tile(4096)
                for rep in range(3):
                    c.reset()
                    c.accumulate(1 << 30, 12288)
                out[tag] = c.h_download(raw=True)
        except Exception as e:  # reported below
            errs.append(repr(e))

    jobs = [("d64", dimer, scp3_amd.FP64), ("hmx", hexamer, scp3_amd.MIXED), ("h64", hexamer, scp3_amd.FP64)]
    th = [threading.Thread(target=work, args=j) for j in jobs]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for tag, sysd, mode in jobs:
        work(tag + "_alone", sysd, mode)
        np.testing.assert_array_equal(out[tag], out[tag + "_alone"])


def test_set_system_revalidates_selected_modes(ctx, dimer):
    """A mode list keeps the caller's order (SCP3 -m "7,0-3"): a new system
    keeps the basis only while every selected mode exists, and a system
    change clears the running sum."""
    setup_system(ctx, dimer, tile=1024)
    ctx.set_basis(np.array([7, 0, 1, 2, 3], dtype=np.int32), 5, 0)
    n = ctx.basis_size()
    ctx.reset()
    ctx.accumulate(1 << 30, 512)
    assert ctx.points_done() == 512
    ctx.set_system(dimer["x0"], dimer["omega"], dimer["MUa"])
    assert ctx.basis_size() == n and ctx.points_done() == 0
    ctx.set_system(dimer["x0"], dimer["omega"][:6], dimer["MUa"][:, :6])
    assert ctx.basis_size() == 0
    with pytest.raises(scp3_amd.Scp3Error):
        ctx.accumulate(1 << 30, 64)


def test_scrambled_points(ctx, dimer):
    """scp3g_set_scramble: the points are L x + e of the reference's Sobol
    integers (the oracle's restatement, checked through the mapped
    Cartesians) for seed 0 and another seed, the two seeds differ, and
    switching scrambling off restores the reference sequence."""
    setup_system(ctx, dimer, tile=1024)
    P, skip = 700, 1 << 30
    M0 = dimer["omega"].size
    X = {}
    for seed in (0, 0x5eed):
        print("scrambled points seed", seed, flush=True)
        ctx.set_scramble(seed)
        X[seed] = ctx.points_coords(skip, P)
        y = oracle.sobol_normals(M0, skip, P, scramble=seed)
        ref = scp3_amd.BOHR * (dimer["x0"][None, :] + y @ dimer["MUa"].T)
        assert np.abs(X[seed] - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(X[0x5eed] - X[0]).max() > 1e-3
    ctx.set_scramble(None)
    X0 = ctx.points_coords(skip, P)
    y0 = oracle.sobol_normals(M0, skip, P)
    ref0 = scp3_amd.BOHR * (dimer["x0"][None, :] + y0 @ dimer["MUa"].T)
    assert np.abs(X0 - ref0).max() < 1e-11 * np.abs(ref0).max()
    assert np.abs(X0 - X[0]).max() > 1e-3
