"""FP64 CPU oracle of the MLP surrogate (forward, loss, gradient).

TEST INFRASTRUCTURE ONLY: checks the device forward/backward kernels of
scp3_amd/csrc/mlp.hip.  The network has no counterpart in the reference.

Parameter layout (FP32 vector, as the C ABI documents):
    [W0 (h x dpad) | b0 (h) | W1 (h x h) | b1 | W2 (h x h) | b2 | w3 (h) | b3]
with W row-major [out][in]; inputs x = sqrt(2) y padded with zeros to dpad.
Model: a_l = silu(W_l a_{l-1} + b_l) for l = 0, 1, 2; f = w3 . a_2 + b3.
Loss: mean over the batch of (f - t)^2.
"""
import numpy as np


def unpack(P, M, h):
    dpad = (M + 63) // 64 * 64
    P = np.asarray(P, dtype=np.float64)
    o = 0
    out = {}
    ins = [dpad, h, h]
    for l in range(3):
        out["W%d" % l] = P[o:o + h * ins[l]].reshape(h, ins[l])
        o += h * ins[l]
        out["b%d" % l] = P[o:o + h]
        o += h
    out["w3"] = P[o:o + h]
    o += h
    out["b3"] = P[o]
    o += 1
    assert o == P.size, (o, P.size)
    return out, dpad


def inputs(y, M, dpad):
    x = np.zeros((y.shape[0], dpad))
    x[:, :M] = np.sqrt(2.0) * y
    return x


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(P, y, M, h):
    W, dpad = unpack(P, M, h)
    a = inputs(y, M, dpad)
    zs, acts = [], [a]
    for l in range(3):
        z = a @ W["W%d" % l].T + W["b%d" % l]
        a = silu(z)
        zs.append(z)
        acts.append(a)
    f = a @ W["w3"] + W["b3"]
    return f, zs, acts


def loss_grad(P, y, t, M, h):
    """Mean squared error and its gradient (flat, same layout as P)."""
    W, dpad = unpack(P, M, h)
    f, zs, acts = forward(P, y, M, h)
    B = y.shape[0]
    r = f - t
    loss = np.mean(r * r)
    dout = 2.0 * r / B
    g = {}
    g["w3"] = acts[3].T @ dout
    g["b3"] = dout.sum()
    delta = np.outer(dout, W["w3"]) * dsilu(zs[2])
    for l in (2, 1, 0):
        g["W%d" % l] = delta.T @ acts[l]
        g["b%d" % l] = delta.sum(0)
        if l > 0:
            delta = (delta @ W["W%d" % l]) * dsilu(zs[l - 1])
    flat = []
    for l in range(3):
        flat += [g["W%d" % l].ravel(), g["b%d" % l]]
    flat += [g["w3"], np.array([g["b3"]])]
    return loss, np.concatenate(flat)


def init_params(M, h, seed):
    """Random parameters for tests (not the device initialiser)."""
    rng = np.random.default_rng(seed)
    dpad = (M + 63) // 64 * 64
    parts = []
    ins = [dpad, h, h]
    for l in range(3):
        fan = M if l == 0 else h
        W = rng.uniform(-1, 1, size=(h, ins[l])) * np.sqrt(6.0 / (fan + h))
        if l == 0:
            W[:, M:] = 0.0
        parts += [W.ravel(), 0.1 * rng.normal(size=h)]
    parts += [rng.uniform(-1, 1, size=h) * np.sqrt(6.0 / (h + 1)), np.array([0.05])]
    return np.concatenate(parts).astype(np.float32)
