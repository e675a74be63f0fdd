"""CPU checks of the gradient (Stein-identity) estimator oracle.

The oracle (oracle/grad_oracle.c) evaluates one rule for every element; the
reference spells each element family out by hand in TripleExcitations.h and
places them in ScaledMatrixElements.cpp:100-725.  These tests pin the oracle
by (1) known-answer values of the reference's own formulas, retyped below
with their file:line, at random points; (2) Gauss-Hermite exactness: for a
polynomial potential the Stein form of every element integrates exactly to
the same matrix as the -G estimator B^T diag(V) B; (3) the reference's
placement differing from the rule only in the five documented families;
(4) q-TIP4P/F gradients against the reference build and finite differences.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O

S2, S3, S6 = math.sqrt(2.0), math.sqrt(3.0), math.sqrt(6.0)


def basis_index(nsel, n2, n3):
    """{sorted ((mode, degree), ...): row} in get_bra order (ScaledMatrixElements.cpp:870-930)."""
    rows = [()]
    rows += [((k, 1),) for k in range(nsel)]
    rows += [((k, 2),) for k in range(n2)]
    rows += [((k, 1), (l, 1)) for k, l in itertools.combinations(range(n2), 2)]
    rows += [((k, 3),) for k in range(n3)]
    for k, l in itertools.combinations(range(n3), 2):
        rows += [((k, 2), (l, 1)), ((k, 1), (l, 2))]
    rows += [((k, 1), (l, 1), (j, 1)) for k, l, j in itertools.combinations(range(n3), 3)]
    return {r: i for i, r in enumerate(rows)}, rows


def one_point(seed, nsel=4):
    rng = np.random.default_rng(seed)
    q = rng.uniform(-1.6, 1.6, nsel)
    vq = rng.uniform(-0.5, 0.5, nsel)
    V = rng.uniform(-0.5, 0.5)
    return q, V, vq


def elem(M, ix, a, b):
    ra, rb = ix[tuple(sorted(a))], ix[tuple(sorted(b))]
    return M[min(ra, rb), max(ra, rb)]


@pytest.mark.parametrize("seed", range(4))
def test_reference_formulas_known_answers(seed):
    """Per-point values of the reference's hand-expanded elements."""
    q, V, vq = one_point(seed)
    ix, _ = basis_index(4, 4, 4)
    M0 = O.grad_accumulate(q[None], [V], vq[None], 4, 4, 4, form=0)
    M1 = O.grad_accumulate(q[None], [V], vq[None], 4, 4, 4, form=1)
    k, l, m, n = 0, 1, 2, 3
    qk, ql, qm, qn = q[[k, l, m, n]]
    Vk, Vl, Vm, Vn = vq[[k, l, m, n]]
    want = {
        # M(0,0) += sum(V), ScaledMatrixElements.cpp:717
        ((), ()): V,
        # f_k3, TripleExcitations.h:27-29
        ((), ((k, 3),)): (4 * V * qk + Vk * (-3 + 2 * qk * qk)) / (2 * S3),
        # m1_f_k2(m, k), :106-108
        (((m, 1),), ((k, 2),)): (qk * qm * Vk + Vm * (-0.5 + qk * qk)) / 2,
        # m1_f_m1, :606-608
        (((m, 1),), ((m, 1),)): V + qm * Vm,
        # m2_f_m2, :692-694
        (((m, 2),), ((m, 2),)): qm * Vm * (-1 + qm * qm) + V * (-0.5 + 3 * qm * qm),
        # m1_n1_f_m1_n1, :1000-1002
        (((m, 1), (n, 1)), ((m, 1), (n, 1))):
            qm * (V * qm + qn * (qn * Vm + qm * Vn)) + V * qn * qn,
        # m3_f_k2_l1(m, k, l), :338-340 (row h3(q_m), column h2(q_k) h1(q_l))
        (((m, 3),), ((k, 2), (l, 1))):
            (4 * V * ql * qm * (-1 + 2 * qk * qk)
             + (2 * qk * ql * qm * Vk + (qm * Vl + ql * Vm) * (-1 + 2 * qk * qk)) * (-3 + 2 * qm * qm))
            / (6 * S3),
    }
    for (a, b), w in want.items():
        assert elem(M0, ix, a, b) == pytest.approx(w, rel=1e-12, abs=1e-14), (a, b)
        assert elem(M1, ix, a, b) == pytest.approx(w, rel=1e-12, abs=1e-14), (a, b)
    # m1_n1_p1_f_k1_l1_j1 (:595-597), 6 distinct modes: needs nsel >= 6
    q6, V6, v6 = one_point(seed + 10, nsel=6)
    ix6, _ = basis_index(6, 6, 6)
    M6 = O.grad_accumulate(q6[None], [V6], v6[None], 6, 6, 6, form=0)
    qm_, qn_, qp_, qk_, ql_, qj_ = q6
    Vm_, Vn_, Vp_, Vk_, Vl_, Vj_ = v6
    w = 2.0 * (qk_ * ql_ * qm_ * qn_ * qp_ * Vj_ + qj_ * ql_ * qm_ * qn_ * qp_ * Vk_
               + qj_ * qk_ * (qm_ * qn_ * qp_ * Vl_ + ql_ * qn_ * qp_ * Vm_ + ql_ * qm_ * qp_ * Vn_
                              + ql_ * qm_ * qn_ * Vp_)) / 3
    got = elem(M6, ix6, ((0, 1), (1, 1), (2, 1)), ((3, 1), (4, 1), (5, 1)))
    assert got == pytest.approx(w, rel=1e-12)
    # m2_n1_f_m1_l1_j1(m, n, l, j), :969-971: bra h2(q_m) h1(q_n), ket h1(q_m) h1(q_l) h1(q_j)
    # with m=0, n=1, l=2, j=3 (rows k2l (0,1) and klj (0,2,3))
    qm_, qn_, ql_, qj_ = q6[:4]
    Vm_, Vn_, Vl_, Vj_ = v6[:4]
    w = (4.0 * V6 * qj_ * ql_ * qm_ * qn_
         + (ql_ * qm_ * qn_ * Vj_ + qj_ * (qm_ * qn_ * Vl_ + ql_ * qn_ * Vm_ + ql_ * qm_ * Vn_))
         * (-1.0 + 2.0 * qm_ * qm_)) / (2 * S2)
    got = elem(M6, ix6, ((0, 2), (1, 1)), ((0, 1), (2, 1), (3, 1)))
    assert got == pytest.approx(w, rel=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_reference_placements(seed):
    """The five element families where the reference's placement departs from the rule."""
    q, V, vq = one_point(seed)
    ix, _ = basis_index(4, 4, 4)
    M0 = O.grad_accumulate(q[None], [V], vq[None], 4, 4, 4, form=0)
    M1 = O.grad_accumulate(q[None], [V], vq[None], 4, 4, 4, form=1)
    k, l = 1, 3
    qk, ql = q[k], q[l]
    Vk, Vl = vq[k], vq[l]
    pair = ((k, 1), (l, 1))
    # the rule at row h1(q_k) is m1_f_m1_l1 (TripleExcitations.h:640-642)
    assert elem(M1, ix, ((k, 1),), pair) == pytest.approx(
        (V * ql + qk * (qk * Vl + ql * Vk)) / S2, rel=1e-12)
    # ... the reference places m2_f_m1_l1(k, l) there (ScaledMatrixElements.cpp:195-196,
    # formula TripleExcitations.h:717-719)
    m2_f_m1_l1 = (4 * V * ql * qk + (qk * Vl + ql * Vk) * (-1 + 2 * qk * qk)) / (2 * S2)
    assert elem(M0, ix, ((k, 1),), pair) == pytest.approx(m2_f_m1_l1, rel=1e-12)
    assert elem(M1, ix, ((k, 2),), pair) == pytest.approx(m2_f_m1_l1, rel=1e-12)
    # row h2(q_k) keeps m2_f_k1_l1(m, k, l) at m = k (ScaledMatrixElements.cpp:192-194,
    # TripleExcitations.h:228-230)
    qm, Vm = qk, Vk
    w = (2 * qk * ql * qm * Vm + (ql * Vk + qk * Vl) * (-1 + 2 * qm * qm)) / (3 * S2)
    assert elem(M0, ix, ((k, 2),), pair) == pytest.approx(w, rel=1e-12)
    # column h3(q_k): row h1 gets m2_f_m3 (ScaledMatrixElements.cpp:264-265,
    # TripleExcitations.h:704-706), the rule's row-h2 value
    x2 = qk * qk
    m2_f_m3 = (Vk * (3 + 4 * x2 * x2 - 8 * x2) + 16 * V * qk * (-1 + x2)) / (2 * S6)
    assert elem(M0, ix, ((k, 1),), ((k, 3),)) == pytest.approx(m2_f_m3, rel=1e-12)
    assert elem(M1, ix, ((k, 2),), ((k, 3),)) == pytest.approx(m2_f_m3, rel=1e-12)
    # row h2 keeps m2_f_k3(m, k) at m = k (:261-263, TripleExcitations.h:213-215)
    w = (4 * V * qk * (-1 + 2 * qm * qm)
         + (-3 + 2 * qk * qk) * (2 * qk * qm * Vm + Vk * (-1 + 2 * qm * qm))) / (4 * S6)
    assert elem(M0, ix, ((k, 2),), ((k, 3),)) == pytest.approx(w, rel=1e-12)
    # diagonal h3: m3_f_m3 (TripleExcitations.h:770-772)
    w = (-3 + 2 * x2) * (qk * Vk * (-3 + 2 * x2) + V * (-3 + 1.0 * x2)) / 6
    assert elem(M0, ix, ((k, 3),), ((k, 3),)) == pytest.approx(w, rel=1e-12)
    # ... whose Stein value has 10 q^2 where the reference has q^2
    w1 = (-3 + 2 * x2) * (qk * Vk * (-3 + 2 * x2) + V * (-3 + 10.0 * x2)) / 6
    assert elem(M1, ix, ((k, 3),), ((k, 3),)) == pytest.approx(w1, rel=1e-12)


def test_reference_form_differs_only_in_five_families():
    nsel, n2, n3 = 5, 5, 4
    rng = np.random.default_rng(7)
    P = 16
    q = rng.normal(size=(P, nsel))
    vq = rng.normal(size=(P, nsel)) * 0.3
    V = rng.normal(size=P) * 0.3
    M0 = O.grad_accumulate(q, V, vq, nsel, n2, n3, 0)
    M1 = O.grad_accumulate(q, V, vq, nsel, n2, n3, 1)
    ix, rows = basis_index(nsel, n2, n3)
    expect = set()
    for k, l in itertools.combinations(range(n2), 2):
        c = ix[((k, 1), (l, 1))]
        for m in (k, l):
            expect.add((ix[((m, 1),)], c))
            expect.add((ix[((m, 2),)], c))
    for k in range(n3):
        c = ix[((k, 3),)]
        expect |= {(ix[((k, 1),)], c), (ix[((k, 2),)], c), (c, c)}
    diff = set(zip(*np.nonzero(np.triu(np.abs(M0 - M1)) > 1e-12)))
    assert diff == expect


def gauss_hermite_points(nmodes, nodes):
    """Tensor Gauss-Hermite rule for the weight exp(-|q|^2)/pi^(d/2)."""
    x, w = np.polynomial.hermite.hermgauss(nodes)
    w = w / math.sqrt(math.pi)
    pts = np.array(list(itertools.product(x, repeat=nmodes)))
    wts = np.prod(np.array(list(itertools.product(w, repeat=nmodes))), axis=1)
    return pts, wts


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 3, 2)])
def test_stein_form_exact_for_polynomial_potential(dims):
    """With Gauss-Hermite nodes both estimators are exact integrals of
    <a|V|b>, so the Stein form must equal B^T diag(V) B to rounding."""
    nsel, n2, n3 = dims
    pts, wts = gauss_hermite_points(nsel, 6)  # exact to degree 11 per mode
    rng = np.random.default_rng(3)
    c1 = rng.normal(size=nsel)
    C2 = rng.normal(size=(nsel, nsel))
    C2 = C2 + C2.T
    c3 = rng.normal(size=nsel) * 0.3
    # V = c0 + c1.q + q^T C2 q / 2 + sum c3 q^3 + q0 q1 q2 (degree 3)
    V = 0.4 + pts @ c1 + 0.5 * np.einsum("pi,ij,pj->p", pts, C2, pts) + (pts ** 3) @ c3
    dV = c1[None, :] + pts @ C2 + 3 * c3[None, :] * pts ** 2
    if nsel >= 3:
        V = V + pts[:, 0] * pts[:, 1] * pts[:, 2]
        dV[:, 0] += pts[:, 1] * pts[:, 2]
        dV[:, 1] += pts[:, 0] * pts[:, 2]
        dV[:, 2] += pts[:, 0] * pts[:, 1]
    Mg = O.grad_accumulate(pts, V * wts, dV * wts[:, None], nsel, n2, n3, 1)
    bra = O.get_bra(pts, nsel, n2, n3)
    n = bra.shape[1]
    Me = np.zeros((n, n), order="F")
    O.add_epot(bra, V * wts, Me)
    U = np.triu(np.ones((n, n), bool))
    assert np.max(np.abs((Mg - Me)[U])) < 1e-12 * max(1.0, np.max(np.abs(Me)))
    # the reference placement is not exact on its five families
    M0 = O.grad_accumulate(pts, V * wts, dV * wts[:, None], nsel, n2, n3, 0)
    assert np.max(np.abs((M0 - Me)[U])) > 1e-6


def test_qtip4pf_gradient_finite_differences(hexamer):
    rng = np.random.default_rng(5)
    x = hexamer["x0"] * 0.52917721092 + rng.normal(size=hexamer["x0"].size) * 0.02
    e, g = O.qtip4pf_grad(x)
    assert e == pytest.approx(O.qtip4pf(x), rel=1e-14)
    h = 1e-5
    fd = np.zeros_like(x)
    for i in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (O.qtip4pf(xp) - O.qtip4pf(xm)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-6 * max(1.0, np.max(np.abs(g)))


@pytest.mark.skipif(not O.has_ref(), reason="reference build absent")
def test_qtip4pf_gradient_matches_reference_build(hexamer):
    import ctypes
    rng = np.random.default_rng(11)
    R = O.ref()
    sysd = hexamer
    for t in range(4):
        x = sysd["x0"] * 0.52917721092 + rng.normal(size=sysd["x0"].size) * 0.05
        x = np.ascontiguousarray(x)
        g_ref = np.zeros_like(x)
        e_ref = R.ref_qtip4pf_grad(x.size // 9, ctypes.c_void_p(x.ctypes.data),
                                   ctypes.c_void_p(g_ref.ctypes.data))
        e, g = O.qtip4pf_grad(x)
        assert e == pytest.approx(e_ref, rel=1e-13, abs=1e-12)
        np.testing.assert_allclose(g, g_ref, rtol=1e-12, atol=1e-11)
