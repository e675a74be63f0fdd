"""The large workloads through the C ABI: the 20825-state hexamer under
TTM3-F (the -G estimator in FP64 and mixed form, and the control variate)
and the 174-mode, 15400-state 20-water cluster.

Bars: FP64 H within 1e-9 of max|H| of the oracle on sampled elements (all
points, ragged counts over several tiles); the mixed form against the FP64
form on the whole upper triangle (relative to max|H|) and its lowest levels
within 0.5 cm^-1 of the FP64 form's on the same points; the control
variate with the same points in both terms equal to the plain estimator to
1e-12; the one-product surrogate term against the hi.hi product formed on
the host.  The oracle evaluates only sampled elements at these sizes.
"""
import numpy as np
import pytest

import scp3_amd
from oracle import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx(gpu_ctx_factory):
    return gpu_ctx_factory(0)


def _prep(name, n2, n3):
    from scp3_amd import system
    return system.prepare(system.input_path(name), nmodes2=n2, nmodes3=n3)


def _setup(c, sysd, pes, mode, tile=0):
    c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
    c.set_scramble(None)
    c.set_basis(sysd["modes"], sysd["nmodes2"], sysd["nmodes3"])
    c.set_pes(pes)
    c.set_precision(mode)
    c.set_tile(tile)
    c.set_mixed_kernel(1)
    c.set_estimator(scp3_amd.EST_ENERGY)


def _sample_elements(n, rng, m=3000):
    ii = rng.integers(0, n, m)
    jj = rng.integers(0, n, m)
    d = rng.integers(0, n, 300)
    ii = np.concatenate([ii, d, np.arange(n - 40, n), rng.integers(0, n, 40), [0, 0]])
    jj = np.concatenate([jj, d, np.full(40, n - 1), np.arange(n - 40, n), [0, n - 1]])
    return np.minimum(ii, jj), np.maximum(ii, jj)


def _oracle_elements(sysd, pes, skip, P, ii, jj, chunk=4096):
    """sum_p bV_p bra_p[ii] bra_p[jj] over Sobol seeds skip.. (the oracle's
    serial potential loop and get_bra), for the sampled (ii, jj)."""
    fp, _ = oracle.pes_pointer(pes, reference=False)
    M0 = sysd["omega"].size
    ref = np.zeros(ii.size)
    for p0 in range(0, P, chunk):
        m = min(chunk, P - p0)
        y = oracle.sobol_normals(M0, skip + p0, m)
        bV = oracle.block_weights(sysd, fp, y)
        bra = oracle.block_bra(sysd, y)
        ref += np.einsum("pk,p,pk->k", bra[:, ii], bV, bra[:, jj])
        del bra
    return ref


def _run_forms(ctx, sysd, pes, skip, P, k):
    out = {}
    for mode in (scp3_amd.FP64, scp3_amd.MIXED):
        print("workload %s n2=%d n3=%d precision %d P=%d" % (pes, sysd["nmodes2"], sysd["nmodes3"], mode, P),
              flush=True)
        _setup(ctx, sysd, pes, mode)
        ctx.reset()
        ctx.accumulate(skip, P)
        out[mode] = (ctx.h_download(raw=True), ctx.lowest_levels(k, 1e-8))
    return out


def test_c3_hexamer_ttm3f_20825_states(ctx):
    """-G with TTM3-F at 20825 states (total quanta <= 3 over 48 modes),
    2^17 + 123 points over five 32768-point tiles."""
    sysd = _prep("hexamer_ttm3f", 48, 48)
    skip, P = 1 << 30, (1 << 17) + 123
    out = _run_forms(ctx, sysd, "ttm3f", skip, P, 10)
    (H64, ev64), (Hmx, evmx) = out[scp3_amd.FP64], out[scp3_amd.MIXED]
    n = H64.shape[0]
    assert n == 20825
    rng = np.random.default_rng(31)
    ii, jj = _sample_elements(n, rng)
    ref = _oracle_elements(sysd, "ttm3f", skip, P, ii, jj)
    err = np.abs(H64[ii, jj] - ref).max() / np.abs(ref).max()
    print("c3 fp64 vs oracle on %d elements: %.2e" % (ii.size, err))
    assert err < 1e-9
    dmx = np.abs(Hmx - H64).max() / np.abs(H64).max()
    dlev = np.abs(evmx - ev64).max() * scp3_amd.AUTOCM
    print("c3 mixed vs fp64: max|dH|/max|H| %.2e, max|dlevel| %.4f cm^-1" % (dmx, dlev))
    assert dmx < 3e-4 and dlev < 0.5
    for H in (H64, Hmx):
        assert not np.any(np.tril(H, -1))


def test_c4_control_variate_20825_states(ctx):
    """The control variate at 20825 states with a 48-256-256-256-1 network
    trained on TTM3-F pairs: (a) the same points in both terms give the
    plain FP64 H to 1e-12; (b) the mixed form's one-product surrogate term
    equals the hi.hi product of X = sqrt|w_nn| B formed on the host from the
    network's own V_nn (scp3g_mlp_eval), on sampled elements."""
    sysd = _prep("hexamer_ttm3f", 48, 48)
    _setup(ctx, sysd, "ttm3f", scp3_amd.FP64)
    ctx.mlp_init(256, seed=0)
    ctx.mlp_sample_data(1 << 40, 1 << 16)
    loss = ctx.mlp_train(400, 4096, 1e-3)
    print("c4 surrogate loss %.4f -> %.4f" % (loss[:10].mean(), loss[-20:].mean()), flush=True)
    assert loss[-20:].mean() < 0.5 * loss[:10].mean()
    skip, P = 1 << 30, (1 << 14) + 77
    ctx.reset()
    ctx.accumulate(skip, P)
    H = ctx.h_download(raw=True)
    ctx.reset()
    ctx.accumulate_cv(skip, P, skip, P)
    Hcv = ctx.h_download(raw=True) * P
    err = np.abs(Hcv - H).max() / np.abs(H).max()
    print("c4 identity: %.2e" % err, flush=True)
    assert err < 1e-12
    del H, Hcv

    # (b) the one-product surrogate term on the mixed form
    ctx.set_precision(scp3_amd.MIXED)
    ctx.set_cv_terms(1)
    P2 = 3 * 4096 + 55
    ctx.reset()
    ctx.accumulate_surrogate(skip, P2)
    Hs = ctx.h_download(raw=True)
    ctx.set_cv_terms(3)
    y = oracle.sobol_normals(48, skip, P2)
    ld = (P2 + 255) // 256 * 256
    ydev = np.zeros((48, ld))
    ydev[:, :P2] = y.T
    yd = torch.from_numpy(ydev).cuda()
    Vd = torch.zeros(P2, dtype=torch.float64, device="cuda")
    import ctypes
    torch.cuda.synchronize()
    ctx.check(ctx.L.scp3g_mlp_eval(ctx.handle, ctypes.c_void_p(yd.data_ptr()), ld, P2,
                                   ctypes.c_void_p(Vd.data_ptr())))
    ctx.sync()
    w = Vd.cpu().numpy() - 0.5 * (y * sysd["omega"] * y).sum(1)
    n = Hs.shape[0]
    rng = np.random.default_rng(5)
    ii, jj = _sample_elements(n, rng, 2000)
    bra = oracle.block_bra(sysd, y)
    sw = np.sqrt(np.abs(w)).astype(np.float32)
    hi = {}
    for key, idx in (("i", ii), ("j", jj)):
        x = sw[:, None] * bra[:, idx].astype(np.float32)
        u = x.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        hi[key] = u.astype(np.uint32).view(np.float32).astype(np.float64)
    H1 = np.einsum("pk,p,pk->k", hi["i"], np.sign(w), hi["j"])
    H64 = np.einsum("pk,p,pk->k", bra[:, ii], w, bra[:, jj])
    got = Hs[ii, jj]
    err1 = np.abs(got - H1).max()
    model = np.abs(H1 - H64).max()
    print("c4 one-product term: |H-H(hi.hi)| %.3e, |H(hi.hi)-H64| %.3e, max|H| %.3e"
          % (err1, model, np.abs(H64).max()))
    assert err1 < 0.05 * model


def test_c5_w20_ttm3f_15400_states(ctx):
    """The 20-water cluster under TTM3-F with all 174 modes and doubles
    (15400 states): FP64 H against the oracle on sampled elements, the mixed
    form against the FP64 form, 2^14 + 33 points over one ragged tile."""
    sysd = _prep("w20_ttm3f", 174, 0)
    skip, P = 1 << 30, (1 << 14) + 33
    out = _run_forms(ctx, sysd, "ttm3f", skip, P, 20)
    (H64, ev64), (Hmx, evmx) = out[scp3_amd.FP64], out[scp3_amd.MIXED]
    n = H64.shape[0]
    assert n == 15400
    rng = np.random.default_rng(17)
    ii, jj = _sample_elements(n, rng)
    ref = _oracle_elements(sysd, "ttm3f", skip, P, ii, jj)
    err = np.abs(H64[ii, jj] - ref).max() / np.abs(ref).max()
    print("c5 fp64 vs oracle on %d elements: %.2e" % (ii.size, err))
    assert err < 1e-9
    dmx = np.abs(Hmx - H64).max() / np.abs(H64).max()
    dlev = np.abs(evmx - ev64).max() * scp3_amd.AUTOCM
    print("c5 mixed vs fp64: max|dH|/max|H| %.2e, max|dlevel| %.4f cm^-1" % (dmx, dlev))
    assert dmx < 3e-4 and dlev < 0.5
