"""MLP surrogate and control variate.

CPU: the FP64 oracle's gradient against central finite differences.
GPU: the device gradient (BF16 weights/activations, FP32 accumulation)
against the FP64 oracle at the same parameters; training lowers the loss;
the control-variate estimator is unbiased (it reproduces the plain FP64 H
up to QMC noise) and, with a trained surrogate, has a smaller error than the
plain estimator at the same number of true-potential evaluations.
"""
import numpy as np
import pytest

from oracle import mlp_oracle


def test_oracle_gradient_is_the_derivative():
    M, h = 12, 64
    P = mlp_oracle.init_params(M, h, 3).astype(np.float64)
    rng = np.random.default_rng(1)
    y = rng.normal(scale=0.7, size=(16, M))
    t = rng.normal(size=16)
    loss, g = mlp_oracle.loss_grad(P, y, t, M, h)
    idx = rng.choice(P.size, 60, replace=False)
    dpad = 64
    for i in idx:
        # padded input columns carry no gradient information
        if i < h * dpad and (i % dpad) >= M:
            continue
        e = np.zeros_like(P)
        e[i] = 1e-6
        fd = (mlp_oracle.loss_grad(P + e, y, t, M, h)[0] - mlp_oracle.loss_grad(P - e, y, t, M, h)[0]) / 2e-6
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


@pytest.mark.gpu
@pytest.mark.parametrize("name,M,h", [("hexamer_qtip4pf", 48, 256), ("w20_ttm3f", 174, 512)])
def test_device_gradient_matches_oracle(gpu_ctx_factory, name, M, h):
    """The device loss gradient against the FP64 oracle at the networks of
    the hexamer (48-256-256-256-1) and the 20-water cluster (174-512-512-
    512-1, inputs padded to 192)."""
    from scp3_amd import system
    sysd = system.prepare(system.input_path(name), nmodes2=1, nmodes3=0)
    assert sysd["omega"].size == M
    print("mlp gradient case %s %d-%d" % (name, M, h), flush=True)
    ctx = gpu_ctx_factory(0)
    ctx.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
    ctx.mlp_init(h, seed=11)
    P = mlp_oracle.init_params(M, h, 5)
    ctx.mlp_set_params(P, np.array([0.0, 0.0, 1.0]))
    rng = np.random.default_rng(2)
    B = 512
    y = rng.normal(scale=np.sqrt(0.5), size=(B, M))
    t = rng.normal(size=B).astype(np.float32)
    ctx.mlp_load_data(y, t)
    g_dev, loss_dev = ctx.mlp_grad(B)
    # the oracle sees the same BF16-rounded layer weights the device multiplies with
    Pb = P.copy()
    u = Pb.view(np.uint32).copy()
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    rb = u.view(np.float32)
    W, dpad = mlp_oracle.unpack(P, M, h)
    assert dpad == (M + 63) // 64 * 64
    o = 0
    for l in range(3):
        n_in = dpad if l == 0 else h
        Pb[o:o + h * n_in] = rb[o:o + h * n_in]
        o += h * n_in + h
    assert o == h * dpad + h + 2 * (h * h + h)
    loss_ref, g_ref = mlp_oracle.loss_grad(Pb.astype(np.float64), y, t.astype(np.float64), M, h)
    rel = np.linalg.norm(g_dev - g_ref) / np.linalg.norm(g_ref)
    cos = g_dev.dot(g_ref) / (np.linalg.norm(g_dev) * np.linalg.norm(g_ref))
    print("mlp gradient: rel.err %.3e  cos %.6f  loss %.5f vs %.5f" % (rel, cos, loss_dev, loss_ref))
    # BF16 operands (8-bit mantissa) in three layers: a few per cent
    assert rel < 5e-2 and cos > 0.998
    assert abs(loss_dev - loss_ref) < 2e-2 * abs(loss_ref)
    del Pb, W, dpad


@pytest.mark.gpu
def test_adam_steps_match_restatement(gpu_ctx_factory, hexamer):
    """Two Adam steps (Kingma & Ba 2015; beta1 0.9, beta2 0.999, eps 1e-8,
    bias-corrected moments) on the FP32 master parameters, against a numpy
    FP32 restatement fed the device's own gradients (scp3g_mlp_grad at the
    same parameters and batch; the training set is one batch, so both steps
    see it).  The first step moves every parameter with a gradient well
    above eps by lr: the bias correction makes m/sqrt(v) = sign(g)."""
    ctx = gpu_ctx_factory(0)
    ctx.set_system(hexamer["x0"], hexamer["omega"], hexamer["MUa"])
    M, h, B, lr = 48, 64, 256, 1e-3
    ctx.mlp_init(h, seed=4)
    P0 = mlp_oracle.init_params(M, h, 9)
    ctx.mlp_set_params(P0, np.array([0.0, 0.0, 1.0]))
    rng = np.random.default_rng(8)
    ctx.mlp_load_data(rng.normal(scale=np.sqrt(0.5), size=(B, M)), rng.normal(size=B).astype(np.float32))
    f32 = np.float32
    b1, b2, eps = f32(0.9), f32(0.999), f32(1e-8)
    m = np.zeros_like(P0)
    v = np.zeros_like(P0)
    P = ctx.mlp_params()[0]
    np.testing.assert_array_equal(P, P0)
    for step in (1, 2):
        g, _ = ctx.mlp_grad(B)
        ctx.mlp_train(1, B, lr)
        P_dev = ctx.mlp_params()[0]
        m = b1 * m + (f32(1) - b1) * g
        v = b2 * v + (f32(1) - b2) * g * g
        c1 = f32(1) - b1 ** step
        c2 = f32(1) - b2 ** step
        P = P - f32(lr) * (m / c1) / (np.sqrt(v / c2) + eps)
        np.testing.assert_allclose(P_dev, P, rtol=0, atol=1e-3 * lr + 4e-7 * np.abs(P).max())
        if step == 1:
            big = np.abs(g) > 1e-4
            assert big.sum() > 1000
            np.testing.assert_allclose(np.abs(P_dev - P0)[big], lr, rtol=2e-2)
        P = P_dev.copy()
    # the update also rewrites the BF16 copies W and W^T that the forward and
    # backward GEMMs read: the gradient through them equals the gradient after
    # the copies are re-derived from the FP32 parameters by mlp_set_params
    g_upd, l_upd = ctx.mlp_grad(B)
    ctx.mlp_set_params(P, np.array([0.0, 0.0, 1.0]))
    g_set, l_set = ctx.mlp_grad(B)
    np.testing.assert_array_equal(g_upd, g_set)
    assert l_upd == l_set


@pytest.mark.gpu
def test_training_lowers_the_loss(gpu_ctx_factory, hexamer):
    ctx = gpu_ctx_factory(0)
    ctx.set_system(hexamer["x0"], hexamer["omega"], hexamer["MUa"])
    ctx.set_pes("qtip4pf")
    ctx.mlp_init(256, seed=0)
    ctx.mlp_sample_data(1 << 30, 1 << 14)
    loss = ctx.mlp_train(600, 1024, 2e-3)
    print("mlp loss: first %.4f  last %.5f" % (loss[:10].mean(), loss[-50:].mean()))
    assert np.all(np.isfinite(loss))
    assert loss[-50:].mean() < 0.2 * loss[:10].mean()


def _small_basis(ctx, hexamer, mode):
    import scp3_amd
    ctx.set_system(hexamer["x0"], hexamer["omega"], hexamer["MUa"])
    ctx.set_basis(np.arange(48, dtype=np.int32), 12, 0)   # 127 states
    ctx.set_pes("qtip4pf")
    ctx.set_precision(mode)
    ctx.set_tile(4096)


@pytest.mark.gpu
def test_control_variate_identity(gpu_ctx_factory, hexamer):
    """With the same points in both terms the control-variate estimator is
    the plain estimator: sum w_nn + sum (w - w_nn) = sum w."""
    import scp3_amd
    ctx = gpu_ctx_factory(0)
    _small_basis(ctx, hexamer, scp3_amd.FP64)
    ctx.mlp_init(64, seed=3)
    ctx.mlp_sample_data(1 << 30, 4096)
    ctx.mlp_train(50, 1024, 1e-3)
    P = 3 * 4096 + 100
    ctx.reset()
    ctx.accumulate(1 << 30, P)
    H = ctx.h_download()
    ctx.reset()
    ctx.accumulate_cv(1 << 30, P, 1 << 30, P)
    Hcv = ctx.h_download()
    assert np.abs(Hcv - H).max() < 1e-12 * np.abs(H).max()


@pytest.mark.gpu
@pytest.mark.parametrize("P", [1 << 14, 3 * 4096 + 100])
def test_control_variate_one_product_identity(gpu_ctx_factory, hexamer, P):
    """Mixed form with the surrogate term on one BF16 product: the
    correction term subtracts w_nn with that same product, so with the same
    points in both terms sum 1p(w_nn) + sum [3p(bV) - 1p(w_nn)] = sum 3p(bV)
    and the control variate equals the plain mixed estimator.  At P = 4^7
    the weights' scale 1/P is a power of 4 and the split-BF16 operands scale
    exactly: the two agree to 1e-8 of max|H| (5e-10 measured); at other P
    to the three-product form's precision.  A correction taking w_nn
    with three products would leave the one-product rounding of the
    surrogate term, measured here as the surrogate term with one against
    three products."""
    import scp3_amd
    print("one-product control variate case: P=%d" % P, flush=True)
    ctx = gpu_ctx_factory(0)
    _small_basis(ctx, hexamer, scp3_amd.MIXED)
    ctx.set_tile(0)
    ctx.mlp_init(64, seed=3)
    ctx.mlp_sample_data(1 << 30, 4096)
    ctx.mlp_train(50, 1024, 1e-3)
    skip = 1 << 30
    ctx.set_cv_terms(1)
    ctx.reset()
    ctx.accumulate(skip, P)
    H = ctx.h_download()
    ctx.reset()
    ctx.accumulate_cv(skip, P, skip, P)
    Hcv = ctx.h_download()
    ctx.reset()
    ctx.accumulate_surrogate(skip, P)
    S1 = ctx.h_download()
    ctx.set_cv_terms(3)
    ctx.reset()
    ctx.accumulate_surrogate(skip, P)
    S3 = ctx.h_download()
    ctx.set_cv_terms(1)
    err = np.abs(Hcv - H).max()
    model = np.abs(S1 - S3).max()
    print("one-product control variate: |H_cv - H| %.3e, one-product rounding %.3e, max|H| %.3e"
          % (err, model, np.abs(H).max()))
    if P == 1 << 14:
        assert err < 1e-8 * np.abs(H).max()
        assert err < 1e-3 * model
    else:
        assert err < 2e-5 * np.abs(H).max()
        assert err < 0.1 * model


@pytest.mark.gpu
def test_hold_as_offset_gives_control_variate(gpu_ctx_factory, hexamer):
    """The correction term held as a fixed term under a surrogate term grown
    in the running sum: Mout, the levels and the sub-block levels equal the
    control variate's S / P_nn + C / P_t (scp3g_accumulate_cv on the same
    seeds, FP64); a second hold adds to the first; raw downloads and a
    cleared term give the running sum alone."""
    import scp3_amd
    ctx = gpu_ctx_factory(0)
    _small_basis(ctx, hexamer, scp3_amd.FP64)
    ctx.mlp_init(64, seed=5)
    ctx.mlp_sample_data(1 << 40, 4096)
    ctx.mlp_train(50, 1024, 1e-3)
    skip_nn, Pnn, skip_t, Pt = 1 << 30, 2 * 4096 + 300, (1 << 30) + 77, 4096 + 11
    ctx.reset()
    ctx.accumulate_cv(skip_nn, Pnn, skip_t, Pt)
    Mref = ctx.h_download()
    ev_ref = ctx.lowest_levels(5)
    n1 = ctx.sub_basis_size(1)
    e1_ref = ctx.lowest_levels_sub(n1, 1)[0]
    ctx.reset()
    with pytest.raises(scp3_amd.Scp3Error):
        ctx.hold_as_offset()                      # nothing accumulated
    ctx.accumulate_correction(skip_t, Pt)
    C = ctx.h_download(raw=True)
    ctx.hold_as_offset()
    assert ctx.points_done() == 0
    ctx.accumulate_surrogate(skip_nn, Pnn)
    S = np.triu(ctx.h_download(raw=True))         # the running sum alone
    C = np.triu(C)
    M = ctx.h_download()
    scale = np.abs(Mref).max()
    np.testing.assert_allclose(M, Mref, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(ctx.lowest_levels(5), ev_ref, rtol=0, atol=1e-10)
    assert abs(ctx.lowest_levels_sub(n1, 1)[0] - e1_ref) < 1e-10
    D = np.diag(np.diag(Mref - S / Pnn - C / Pt))  # the HO diagonal
    # a second hold adds the running sum to the held term
    ctx.hold_as_offset()
    ctx.accumulate_correction(skip_t, Pt)
    np.testing.assert_allclose(ctx.h_download(), Mref + C / Pt, rtol=0, atol=1e-12 * scale)
    # cleared: the running sum alone
    ctx.clear_offset()
    np.testing.assert_allclose(ctx.h_download(), C / Pt + D, rtol=0, atol=1e-12 * scale)


@pytest.mark.gpu
def test_control_variate_beats_plain_qmc(gpu_ctx_factory, hexamer):
    """At the same number of true-potential points, the trained surrogate as
    control variate gives a smaller error of H and of the levels than plain
    QMC (reference: plain FP64 QMC at 2^20 points)."""
    import scp3_amd
    ctx = gpu_ctx_factory(0)
    _small_basis(ctx, hexamer, scp3_amd.FP64)
    ctx.reset()
    ctx.accumulate(1 << 30, 1 << 20)
    Href = ctx.h_download()
    ev_ref = ctx.lowest_levels(10)
    Pt = 1 << 12
    ctx.reset()
    ctx.accumulate(1 << 31, Pt)          # plain: Pt true points (another stretch of the sequence)
    Hpl = ctx.h_download()
    ev_pl = ctx.lowest_levels(10)
    ctx.mlp_init(256, seed=0)
    ctx.mlp_sample_data(1 << 40, 1 << 15)
    loss = ctx.mlp_train(3000, 1024, 1e-3)
    ctx.reset()
    ctx.accumulate_cv(1 << 31, 1 << 18, 1 << 31, Pt)
    Hcv = ctx.h_download()
    ev_cv = ctx.lowest_levels(10)
    e_pl = np.linalg.norm(Hpl - Href)
    e_cv = np.linalg.norm(Hcv - Href)
    l_pl = np.abs(ev_pl - ev_ref).max() * scp3_amd.AUTOCM
    l_cv = np.abs(ev_cv - ev_ref).max() * scp3_amd.AUTOCM
    print("control variate: loss %.4f, |dH|_F plain %.3e cv %.3e, max|dlevel| plain %.3f cv %.3f cm^-1"
          % (loss[-100:].mean(), e_pl, e_cv, l_pl, l_cv))
    assert e_cv < e_pl
    assert l_cv < l_pl


def _bf16_rne(x32):
    u = x32.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("terms", [1, 3])
def test_surrogate_term_split_bf16_products(gpu_ctx_factory, hexamer, terms):
    """The surrogate term on the mixed form's signed split-BF16 SYRK with a
    constant surrogate w_nn = c < 0 (all sign bits set): against the
    hi.hi-only product computed on the host (terms = 1) or the FP64 product
    (terms = 3), on the same Sobol points."""
    import scp3_amd
    from oracle import oracle
    print("surrogate term case: terms=%d" % terms, flush=True)
    ctx = gpu_ctx_factory(0)
    _small_basis(ctx, hexamer, scp3_amd.MIXED)
    ctx.set_tile(0)
    ctx.set_cv_terms(terms)
    ctx.mlp_init(64, seed=1)
    P0, _ = ctx.mlp_params()
    c = -0.0123
    ctx.mlp_set_params(np.zeros_like(P0), np.array([c, 0.0, 1.0]))   # f = 0, w_nn = c
    skip, P = 1 << 30, 3 * 4096 + 777
    ctx.reset()
    ctx.accumulate_surrogate(skip, P)
    H = ctx.h_download(raw=True)
    y = oracle.sobol_normals(48, skip, P)
    b = oracle.get_bra(y[:, hexamer["modes"]], 48, 12, 0)          # P x n, FP64
    H64 = c * (b.T @ b)
    x = np.float32(np.sqrt(abs(c))) * b.astype(np.float32)
    hi = _bf16_rne(x).astype(np.float64)
    H1 = np.sign(c) * (hi.T @ hi)
    iu = np.triu_indices(b.shape[1])
    err1 = np.abs(H - H1)[iu].max()
    err64 = np.abs(H - H64)[iu].max()
    model = np.abs(H1 - H64)[iu].max()
    print("surrogate term terms=%d: |H-H(hi.hi)| %.3e  |H-H64| %.3e  |H(hi.hi)-H64| %.3e  max|H| %.3e"
          % (terms, err1, err64, model, np.abs(H64).max()))
    if terms == 1:
        # the device computes exactly the hi.hi product (up to FP32 sums and
        # rare one-ulp differences of the FP32 basis values)
        assert err1 < 0.02 * model
    else:
        assert err64 < 1e-5 * np.abs(H64).max()
        assert err64 < 0.02 * model


@pytest.mark.gpu
def test_control_variate_terms_separately(gpu_ctx_factory, hexamer):
    """scp3g_accumulate_surrogate and scp3g_accumulate_correction, each in a
    sum of its own, add up to scp3g_accumulate_cv on the same seeds:
    H_cv = S / P_nn + C / P_t (FP64, raw sums)."""
    import scp3_amd
    ctx = gpu_ctx_factory(0)
    _small_basis(ctx, hexamer, scp3_amd.FP64)
    ctx.mlp_init(64, seed=5)
    ctx.mlp_sample_data(1 << 40, 4096)
    ctx.mlp_train(50, 1024, 1e-3)
    skip_nn, Pnn, skip_t, Pt = 1 << 30, 2 * 4096 + 300, (1 << 30) + 77, 4096 + 11
    ctx.reset()
    ctx.accumulate_surrogate(skip_nn, Pnn)
    S = ctx.h_download(raw=True)
    ctx.reset()
    ctx.accumulate_correction(skip_t, Pt)
    C = ctx.h_download(raw=True)
    assert ctx.points_done() == Pt
    ctx.reset()
    ctx.accumulate_cv(skip_nn, Pnn, skip_t, Pt)
    H = ctx.h_download(raw=True)
    ref = S / Pnn + C / Pt
    assert np.abs(H - ref).max() < 1e-12 * np.abs(ref).max()
    # each term's Mout is its mean plus the same HO diagonal
    ctx.reset()
    ctx.accumulate_correction(skip_t, Pt)
    Mc = ctx.h_download()
    ctx.reset()
    ctx.accumulate_surrogate(skip_nn, Pnn)
    Ms = ctx.h_download()
    np.testing.assert_allclose(Mc - C / Pt, Ms - S / Pnn, rtol=0, atol=1e-12 * np.abs(ref).max())
