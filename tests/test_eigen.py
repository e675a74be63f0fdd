"""Lowest levels: the host Rayleigh-Ritz solver (CPU) and the GPU Davidson
solver behind scp3g_lowest_levels / scp3g_matrix_lowest_levels, which
replace the eigensolver's dsyevr (eigensolver.cpp:20-80) and SCP3's eig_sym
(SCP3.cpp:242,246).  References: LAPACK through numpy/scipy on the same
matrix."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def small_eigh_bin(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("se") / "small_eigh")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "small_eigh_main.cpp"),
                    os.path.join(ROOT, "scp3_amd", "csrc", "small_eigh.cpp"), "-o", exe], check=True)
    return exe


def _run(exe, A, nvec):
    m = A.shape[0]
    inp = "%d %d\n" % (m, nvec) + "\n".join("%.17e" % x for x in A.flatten(order="F"))
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
    v = np.array([float(x) for x in out])
    nv = min(nvec, m)
    return v[:nv], v[nv:].reshape(nv, m).T


def _cases():
    rng = np.random.default_rng(0)
    for m in (1, 2, 3, 17, 64, 128):
        B = rng.normal(size=(m, m))
        yield "random", B + B.T
        yield "diagonal", np.diag(rng.normal(size=m))
        # exact threefold degeneracy
        yield "degenerate", np.diag(np.repeat(np.arange((m + 2) // 3), 3)[:m].astype(float))
        # fourfold near-degeneracy in a rotated basis
        Q, _ = np.linalg.qr(rng.normal(size=(m, m)))
        ev = np.repeat(np.arange((m + 3) // 4), 4)[:m] + 1e-13 * rng.normal(size=m)
        yield "rotated-degenerate", Q @ np.diag(ev) @ Q.T
        yield "zero", np.zeros((m, m))
        # the Davidson projected problem: dominant diagonal, weak coupling
        C = rng.normal(size=(m, m))
        yield "diagonally-dominant", np.diag(np.linspace(0.07, 0.2, m)) + 1e-4 * (C + C.T)


@pytest.mark.parametrize("name,A", list(_cases()), ids=lambda x: x if isinstance(x, str) else str(x.shape[0]))
def test_small_eigh_against_lapack(small_eigh_bin, name, A):
    m = A.shape[0]
    nvec = min(m, 28)
    w, Z = _run(small_eigh_bin, A, nvec)
    scale = max(np.abs(A).max(), 1.0)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(A)[:nvec], rtol=0, atol=1e-13 * scale * max(m, 1))
    assert np.abs(A @ Z - Z * w).max() < 1e-11 * scale
    assert np.abs(Z.T @ Z - np.eye(nvec)).max() < 1e-11


def _dd_matrix(n, seed, degenerate=False):
    """Symmetric, diagonally dominant like Mout: HO-like diagonal (degenerate
    multiplets optional) and a weak dense coupling."""
    rng = np.random.default_rng(seed)
    if degenerate:
        d = 0.05 + 0.004 * np.floor(np.sqrt(np.arange(n)))
    else:
        d = np.sort(0.05 + 0.1 * rng.random(n))
    C = rng.normal(size=(n, n)) * 2e-5
    return np.diag(d) + C + C.T


@pytest.mark.gpu
@pytest.mark.parametrize("n,k,degenerate", [(20, 5, False), (91, 5, False), (300, 10, False),
                                            (1225, 10, True), (2001, 28, False), (4099, 10, True)])
def test_matrix_lowest_levels(gpu_ctx_factory, n, k, degenerate):
    import scipy.linalg
    print("eigen case: n=%d k=%d degenerate=%s" % (n, k, degenerate), flush=True)
    ctx = gpu_ctx_factory(0)
    A = _dd_matrix(n, n, degenerate)
    ev, V = ctx.matrix_lowest_levels(np.asfortranarray(A), k, vectors=True)
    ref = scipy.linalg.eigh(A, eigvals_only=True, subset_by_index=[0, k - 1])
    np.testing.assert_allclose(ev, ref, rtol=0, atol=1e-10)
    # eigenvectors: residual at the Davidson tolerance (1e-9), orthonormal
    R = A @ V - V * ev
    assert np.linalg.norm(R, axis=0).max() < 2e-9
    assert np.abs(V.T @ V - np.eye(k)).max() < 1e-8
