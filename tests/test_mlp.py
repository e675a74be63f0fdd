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
def test_device_gradient_matches_oracle(gpu_ctx_factory, hexamer):
    import scp3_amd
    ctx = gpu_ctx_factory(0)
    ctx.set_system(hexamer["x0"], hexamer["omega"], hexamer["MUa"])
    M, h = 48, 256
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
    nW = 3 * h * 0 + h * 64 + h + 2 * (h * h + h)
    u = Pb.view(np.uint32).copy()
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    rb = u.view(np.float32)
    W, dpad = mlp_oracle.unpack(P, M, h)
    o = 0
    for l in range(3):
        n_in = dpad if l == 0 else h
        Pb[o:o + h * n_in] = rb[o:o + h * n_in]
        o += h * n_in + h
    assert o == nW
    loss_ref, g_ref = mlp_oracle.loss_grad(Pb.astype(np.float64), y, t.astype(np.float64), M, h)
    rel = np.linalg.norm(g_dev - g_ref) / np.linalg.norm(g_ref)
    cos = g_dev.dot(g_ref) / (np.linalg.norm(g_dev) * np.linalg.norm(g_ref))
    print("mlp gradient: rel.err %.3e  cos %.6f  loss %.5f vs %.5f" % (rel, cos, loss_dev, loss_ref))
    # BF16 operands (8-bit mantissa) in three layers: a few per cent
    assert rel < 5e-2 and cos > 0.998
    assert abs(loss_dev - loss_ref) < 2e-2 * abs(loss_ref)
    del Pb, W, dpad


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
