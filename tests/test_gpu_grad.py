"""Gradient (Stein-identity) estimator on the GPU vs the CPU oracle, through the C ABI.

Bars: q-TIP4P/F energies to 1e-12 relative and gradients to 1e-10 of their
largest component against oracle/grad_oracle.c; the raw running sum M
(before division by the points) within 1e-9 of max|M| of the oracle's on
the same Sobol points, for the reference's element placement and for the
Stein form; the lowest levels within 1e-8 hartree of LAPACK on the oracle's
finalised matrix.  The FP64 sums differ from the oracle's only in summation
order (split-K partials, the +-(phi +- s Psi) SYRK pair and the block
corrections).
"""
import numpy as np
import pytest

import scp3_amd
from oracle import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# oracle form numbers: 0 = the reference's placement, 1 = the Stein rule
FORMS = {scp3_amd.EST_GRADIENT: 0, scp3_amd.EST_GRADIENT_STEIN: 1}


def vp(t):
    import ctypes
    torch.cuda.synchronize()
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def ctx(gpu_ctx_factory):
    return gpu_ctx_factory(0)


def setup(c, sysd, est, tile=4096):
    c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
    c.set_basis(sysd["modes"], sysd["nmodes2"], sysd["nmodes3"])
    c.set_pes("qtip4pf")
    c.set_precision(scp3_amd.FP64)
    c.set_tile(tile)
    c.set_estimator(est)


@pytest.mark.parametrize("which", ["dimer", "hexamer"])
def test_qtip4pf_gradient_matches_oracle(ctx, which, dimer, hexamer):
    sysd = dimer if which == "dimer" else hexamer
    print("qtip4pf gradient", which, flush=True)
    rng = np.random.default_rng(2)
    n3 = sysd["x0"].size
    nw = n3 // 9
    P = 1000
    X = sysd["x0"][None, :] * scp3_amd.BOHR + rng.normal(size=(P, n3)) * 0.04
    ld = 1024
    crd = torch.zeros(n3 * ld, dtype=torch.float64, device="cuda")
    crd.view(n3, ld)[:, :P] = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    E = torch.zeros(ld, dtype=torch.float64, device="cuda")
    G = torch.zeros(n3 * ld, dtype=torch.float64, device="cuda")
    ctx.check(ctx.L.scp3g_pes_gradient(ctx.handle, scp3_amd.PES_QTIP4PF, nw, P, vp(crd), ld, vp(E), vp(G), ld))
    ctx.sync()
    Eg = E.cpu().numpy()[:P]
    Gg = G.cpu().numpy().reshape(n3, ld)[:, :P].T
    for i in range(0, P, 37):
        e, g = oracle.qtip4pf_grad(X[i])
        assert Eg[i] == pytest.approx(e, rel=1e-12, abs=1e-11)
        np.testing.assert_allclose(Gg[i], g, rtol=0, atol=1e-10 * np.max(np.abs(g)))


def _check_sums(ctx, sysd, est, seed, npts, tile=4096):
    setup(ctx, sysd, est, tile)
    ctx.accumulate(seed, npts)
    M = ctx.h_download(raw=True)
    ref = oracle.scp3_accumulate_grad(sysd, seed, npts, form=FORMS[est])["M"]
    U = np.triu(np.ones(M.shape, bool))
    scale = np.max(np.abs(ref[U]))
    err = np.max(np.abs((M - ref)[U]))
    print("grad sums", est, M.shape[0], npts, "max err / max", err / scale, flush=True)
    assert err <= 1e-9 * scale
    assert np.all(M[~U] == 0.0)
    return M, ref


@pytest.mark.parametrize("est", [scp3_amd.EST_GRADIENT, scp3_amd.EST_GRADIENT_STEIN])
def test_grad_sums_dimer(ctx, dimer, est):
    # 5000 points: one full 4096-point tile and a ragged second one
    _check_sums(ctx, dimer, est, 1 << 30, 5000)


@pytest.mark.parametrize("est", [scp3_amd.EST_GRADIENT, scp3_amd.EST_GRADIENT_STEIN])
def test_grad_sums_dimer_triples(ctx, est):
    from scp3_amd import system
    sysd = system.prepare(system.input_path("dimer_qtip4pf"), nmodes2=12, nmodes3=7)
    _check_sums(ctx, sysd, est, 12345, 3000)


def test_grad_sums_hexamer(ctx, hexamer):
    _check_sums(ctx, hexamer, scp3_amd.EST_GRADIENT, 1 << 30, 300)


def test_grad_levels_dimer(ctx, dimer):
    import scipy.linalg
    npts = 4096
    setup(ctx, dimer, scp3_amd.EST_GRADIENT)
    ctx.accumulate(1 << 30, npts)
    ev = ctx.lowest_levels(5, tol=1e-10)
    Mout = ctx.h_download()
    ref = oracle.finalize(dimer, oracle.scp3_accumulate_grad(dimer, 1 << 30, npts, form=0)["M"], npts)
    np.testing.assert_allclose(np.triu(Mout), np.triu(ref), rtol=0, atol=1e-12 * np.max(np.abs(ref)))
    w = scipy.linalg.eigh(ref, lower=False, eigvals_only=True, subset_by_index=[0, 4])
    np.testing.assert_allclose(ev, w, rtol=0, atol=1e-8)


def test_grad_normals_branch(ctx, dimer):
    """The -r branch (caller's standard normals) through the gradient estimator."""
    rng = np.random.default_rng(9)
    npts = 700
    Z = rng.normal(size=(npts, dimer["omega"].size))
    setup(ctx, dimer, scp3_amd.EST_GRADIENT)
    ctx.accumulate_normals(Z)
    M = ctx.h_download(raw=True)
    ref = oracle.scp3_accumulate_grad(dimer, 0, npts, form=0, normals=Z)["M"]
    U = np.triu(np.ones(M.shape, bool))
    assert np.max(np.abs((M - ref)[U])) <= 1e-9 * np.max(np.abs(ref[U]))


def test_grad_estimator_rejects_unsupported(ctx, dimer):
    setup(ctx, dimer, scp3_amd.EST_GRADIENT)
    ctx.set_precision(scp3_amd.MIXED)
    ctx.set_mixed_kernel(0)
    with pytest.raises(scp3_amd.Scp3Error, match="LDS-tiled"):
        ctx.accumulate(0, 256)
    ctx.set_mixed_kernel(1)
    ctx.set_precision(scp3_amd.FP64)
    with pytest.raises(scp3_amd.Scp3Error, match="energy-input"):
        ctx.accumulate_energies(0, np.zeros(256))
    ctx.set_estimator(scp3_amd.EST_ENERGY)


@pytest.mark.parametrize("name", ["hexamer_ttm3f", "w20_ttm3f"])
def test_ttm3f_gradient_matches_oracle(ctx, name):
    from scp3_amd import system
    print("ttm3f gradient", name, flush=True)
    sysd = system.prepare(system.input_path(name), nmodes2=0, nmodes3=0)
    rng = np.random.default_rng(3)
    n3 = sysd["x0"].size
    nw = n3 // 9
    P = 200
    X = sysd["x0"][None, :] * scp3_amd.BOHR + rng.normal(size=(P, n3)) * 0.03
    ld = 256
    crd = torch.zeros(n3 * ld, dtype=torch.float64, device="cuda")
    crd.view(n3, ld)[:, :P] = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    E = torch.zeros(ld, dtype=torch.float64, device="cuda")
    G = torch.zeros(n3 * ld, dtype=torch.float64, device="cuda")
    ctx.check(ctx.L.scp3g_pes_gradient(ctx.handle, scp3_amd.PES_TTM3F, nw, P, vp(crd), ld, vp(E), vp(G), ld))
    ctx.sync()
    Eg = E.cpu().numpy()[:P]
    Gg = G.cpu().numpy().reshape(n3, ld)[:, :P].T
    for i in range(0, P, 23):
        e, g = oracle.ttm3f_grad(X[i])
        assert Eg[i] == pytest.approx(e, rel=1e-12, abs=1e-10)
        np.testing.assert_allclose(Gg[i], g, rtol=0, atol=1e-9 * np.max(np.abs(g)))


def test_ttm3f_gradient_across_scratch_chunks(ctx):
    """scp3g_pes_gradient runs TTM3-F configurations in chunks sized by its
    ~1 GiB monomer and pair scratch (17567 configurations for 20 waters:
    8 x (40 nw + 36 nw(nw-1)/2) bytes each).  35634 configurations cross two
    chunk boundaries with a ragged end; configurations on both sides of each
    boundary and spread over the range match the oracle."""
    from scp3_amd import system
    sysd = system.prepare(system.input_path("w20_ttm3f"), nmodes2=0, nmodes3=0)
    n3 = sysd["x0"].size
    nw = n3 // 9
    per_cfg = 8 * (40 * nw + 36 * nw * (nw - 1) // 2)
    chunk = (1 << 30) // per_cfg
    P = 2 * chunk + 500
    print("ttm3f gradient chunks: nw %d, %d configurations, chunk %d" % (nw, P, chunk), flush=True)
    rng = np.random.default_rng(7)
    X = sysd["x0"][None, :] * scp3_amd.BOHR + rng.normal(size=(P, n3)) * 0.03
    crd = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    E = torch.zeros(P, dtype=torch.float64, device="cuda")
    G = torch.zeros(n3 * P, dtype=torch.float64, device="cuda")
    ctx.check(ctx.L.scp3g_pes_gradient(ctx.handle, scp3_amd.PES_TTM3F, nw, P, vp(crd), P, vp(E), vp(G), P))
    ctx.sync()
    Eg = E.cpu().numpy()
    Gg = G.cpu().numpy().reshape(n3, P)
    check = set(range(0, P, 997)) | {P - 1}
    for b in (chunk, 2 * chunk):
        check |= set(range(b - 3, b + 3))
    for i in sorted(check):
        e, g = oracle.ttm3f_grad(X[i])
        assert Eg[i] == pytest.approx(e, rel=1e-12, abs=1e-10), i
        np.testing.assert_allclose(Gg[:, i], g, rtol=0, atol=1e-9 * np.max(np.abs(g)), err_msg=str(i))


def test_grad_sums_hexamer_ttm3f(ctx):
    """The gradient estimator with TTM3-F (the oracle's TTM3-F gradient)."""
    from scp3_amd import system
    sysd = system.prepare(system.input_path("hexamer_ttm3f"), nmodes2=12, nmodes3=6, modes=None)
    setup(ctx, sysd, scp3_amd.EST_GRADIENT)
    ctx.set_pes("ttm3f")
    npts = 600
    ctx.accumulate(1 << 30, npts)
    M = ctx.h_download(raw=True)
    ref = oracle.scp3_accumulate_grad(sysd, 1 << 30, npts, form=0, pes="ttm3f")["M"]
    U = np.triu(np.ones(M.shape, bool))
    assert np.max(np.abs((M - ref)[U])) <= 1e-9 * np.max(np.abs(ref[U]))


@pytest.mark.parametrize("est", [scp3_amd.EST_GRADIENT_STEIN, scp3_amd.EST_GRADIENT])
def test_grad_mixed_matches_fp64(ctx, hexamer, est):
    """The rank-2 term on split-BF16 MFMA against the FP64 form on the same points."""
    P = 1 << 15
    out = {}
    for mode in (scp3_amd.FP64, scp3_amd.MIXED):
        setup(ctx, hexamer, est, tile=16384)
        ctx.set_precision(mode)
        ctx.accumulate(1 << 30, P)
        out[mode] = (ctx.h_download(), ctx.lowest_levels(10))
    (H64, ev64), (Hmx, evmx) = out[scp3_amd.FP64], out[scp3_amd.MIXED]
    dH = np.abs(np.triu(Hmx - H64)).max() / np.abs(np.triu(H64)).max()
    dlev = np.abs(evmx - ev64).max() * scp3_amd.AUTOCM
    print("gradient mixed vs fp64: max|dH|/max|H| %.3e, max|dlevel| %.4f cm^-1" % (dH, dlev), flush=True)
    assert dH < 1e-5
    assert dlev < 0.5
    ctx.set_precision(scp3_amd.FP64)


def test_grad_mixed_with_triples_matches_fp64(ctx):
    """Blocks of more than 64 functions (triples): the per-mode corrections on
    split-BF16 MFMA in the mixed form, against FP64 on the same points."""
    from scp3_amd import system
    sysd = system.prepare(system.input_path("hexamer_qtip4pf"), nmodes2=48, nmodes3=12)
    P = 1 << 14
    out = {}
    for mode in (scp3_amd.FP64, scp3_amd.MIXED):
        setup(ctx, sysd, scp3_amd.EST_GRADIENT_STEIN, tile=8192)
        ctx.set_precision(mode)
        ctx.accumulate(1 << 30, P)
        out[mode] = (ctx.h_download(), ctx.lowest_levels(10))
    (H64, ev64), (Hmx, evmx) = out[scp3_amd.FP64], out[scp3_amd.MIXED]
    dH = np.abs(np.triu(Hmx - H64)).max() / np.abs(np.triu(H64)).max()
    dlev = np.abs(evmx - ev64).max() * scp3_amd.AUTOCM
    print("gradient mixed (triples) vs fp64: max|dH|/max|H| %.3e, max|dlevel| %.4f cm^-1" % (dH, dlev), flush=True)
    assert dH < 1e-5
    assert dlev < 0.5
    ctx.set_precision(scp3_amd.FP64)


def test_grad_mixed_w20_matches_fp64(ctx):
    """174 selected modes (the 20-water cluster, TTM3-F): the rank-2 builder
    with every mode in LDS, mixed against FP64 on the same points."""
    from scp3_amd import system
    sysd = system.prepare(system.input_path("w20_ttm3f"), nmodes2=10, nmodes3=0)
    P = 1 << 13
    out = {}
    for mode in (scp3_amd.FP64, scp3_amd.MIXED):
        setup(ctx, sysd, scp3_amd.EST_GRADIENT_STEIN, tile=4096)
        ctx.set_pes("ttm3f")
        ctx.set_precision(mode)
        ctx.accumulate(1 << 30, P)
        out[mode] = (ctx.h_download(), ctx.lowest_levels(10))
    (H64, ev64), (Hmx, evmx) = out[scp3_amd.FP64], out[scp3_amd.MIXED]
    dH = np.abs(np.triu(Hmx - H64)).max() / np.abs(np.triu(H64)).max()
    dlev = np.abs(evmx - ev64).max() * scp3_amd.AUTOCM
    print("gradient mixed (w20) vs fp64: max|dH|/max|H| %.3e, max|dlevel| %.4f cm^-1" % (dH, dlev), flush=True)
    assert dH < 1e-5
    assert dlev < 0.5
    ctx.set_precision(scp3_amd.FP64)


def test_reference_form_differs_only_in_five_families_large(ctx):
    """80 doubly excited modes (6321 fixed elements, two blocks of fixes): the
    reference's placement equals the Stein form outside the five families."""
    import itertools
    from scp3_amd import system
    sysd = system.prepare(system.input_path("w20_ttm3f"), nmodes2=80, nmodes3=0)
    P = 2048
    M = {}
    for est in (scp3_amd.EST_GRADIENT, scp3_amd.EST_GRADIENT_STEIN):
        setup(ctx, sysd, est, tile=2048)
        ctx.set_pes("ttm3f")
        ctx.accumulate(1 << 30, P)
        M[est] = ctx.h_download(raw=True)
    D = np.abs(np.triu(M[scp3_amd.EST_GRADIENT] - M[scp3_amd.EST_GRADIENT_STEIN]))
    nsel, n2 = sysd["modes"].size, 80
    fam = np.zeros(D.shape, bool)
    c = 1 + nsel + n2
    for k, l in itertools.combinations(range(n2), 2):
        for m in (k, l):
            fam[1 + m, c] = fam[1 + nsel + m, c] = True
        c += 1
    scale = np.abs(M[scp3_amd.EST_GRADIENT_STEIN]).max()
    assert D[~fam].max() <= 1e-12 * scale
    assert np.median(D[fam]) > 1e-9 * scale


@pytest.mark.parametrize("est", [scp3_amd.EST_GRADIENT, scp3_amd.EST_GRADIENT_STEIN])
def test_accumulate_gradients_matches_device_potential(ctx, dimer, est):
    """Energies and gradients from the caller (the oracle's q-TIP4P/F on the
    Cartesians of points_coords) give the device path's sums."""
    npts = 3000
    setup(ctx, dimer, est, tile=2048)
    X = ctx.points_coords(1 << 30, npts)
    E = np.zeros(npts)
    G = np.zeros_like(X)
    for i in range(npts):
        E[i], G[i] = oracle.qtip4pf_grad(X[i])
    ctx.reset()
    ctx.accumulate_gradients(1 << 30, E, G)
    M1 = ctx.h_download(raw=True)
    ctx.reset()
    ctx.accumulate(1 << 30, npts)
    M2 = ctx.h_download(raw=True)
    U = np.triu(np.ones(M1.shape, bool))
    assert np.max(np.abs((M1 - M2)[U])) <= 1e-9 * np.max(np.abs(M2[U]))
    with pytest.raises(scp3_amd.Scp3Error, match="gradient estimator"):
        ctx.set_estimator(scp3_amd.EST_ENERGY)
        ctx.accumulate_gradients(1 << 30, E[:10], G[:10])


def test_stein_mixed_matches_fp64_20825_states(ctx):
    """The full hexamer TTM3-F basis (20825 states): mixed against FP64 Stein
    form on the same 2^15 points, lowest ten levels within 0.5 cm^-1."""
    from scp3_amd import system
    print("stein 20825 states mixed vs fp64", flush=True)
    sysd = system.prepare(system.input_path("hexamer_ttm3f"), nmodes2=48, nmodes3=48)
    P = 1 << 15
    ev = {}
    for mode in (scp3_amd.FP64, scp3_amd.MIXED):
        setup(ctx, sysd, scp3_amd.EST_GRADIENT_STEIN, tile=0)
        ctx.set_pes("ttm3f")
        ctx.set_precision(mode)
        ctx.accumulate(1 << 30, P)
        ev[mode] = ctx.lowest_levels(10, tol=1e-7)
    dlev = np.abs(ev[scp3_amd.MIXED] - ev[scp3_amd.FP64]).max() * scp3_amd.AUTOCM
    print("stein 20825: max|dlevel| %.4f cm^-1" % dlev, flush=True)
    assert dlev < 0.5
    ctx.set_precision(scp3_amd.FP64)


@pytest.mark.parametrize("est", [scp3_amd.EST_GRADIENT, scp3_amd.EST_GRADIENT_STEIN])
def test_grad_sums_mode_subset(ctx, est):
    """A selection of modes (-m): the estimator over the selected modes with
    bV over all of them (SCP3.cpp:219-231)."""
    from scp3_amd import system
    sysd = system.prepare(system.input_path("dimer_qtip4pf"), nmodes2=5, nmodes3=3,
                          modes=[1, 3, 4, 6, 7, 9, 11])
    _check_sums(ctx, sysd, est, 777, 2500)
