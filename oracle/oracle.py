"""ctypes wrapper over the CPU FP64 oracle (liboracle.so, scp3_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by
bench.py's cpu_baseline leg to check and time the product; the product
library never loads it.  `ref()` loads the reference's own CPU sources
compiled by oracle/Makefile into oracle/_ref/libscp3ref.so when present.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "libscp3ref.so")

_orc = None
_ref = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
I = ctypes.c_int
D = ctypes.c_double


def build(ref=True):
    """Compile liboracle.so (and the reference build when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"] + (["ref"] if ref else []), check=True)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = ctypes.CDLL(LIB)
        L.orc_sobol_ints.argtypes = [I, I64, I64, P]
        L.orc_sobol_normals.argtypes = [I, I64, I64, I, P]
        L.orc_sobol_normals_scr.argtypes = [I, I64, I64, I, I, ctypes.c_uint64, P]
        L.orc_sobol_scramble.argtypes = [I, ctypes.c_uint64, I64, P]
        L.orc_sobol_table.argtypes = [I, P]
        L.orc_bsm.argtypes = [I64, P, P]
        L.orc_ho.argtypes = [I, D]
        L.orc_ho.restype = D
        L.orc_basis_size.argtypes = [I, I, I]
        L.orc_basis_size.restype = I64
        L.orc_sub_basis_size.argtypes = [I, I, I, I]
        L.orc_sub_basis_size.restype = I64
        L.orc_get_bra.argtypes = [I, I, I, P, P]
        L.orc_get_bra.restype = I64
        L.orc_add_epot.argtypes = [I64, I64, P, P, P]
        L.orc_add_ho_diagonal.argtypes = [I, I, I, P, I64, P]
        L.orc_qtip4pf.argtypes = [I64, P]
        L.orc_qtip4pf.restype = D
        L.orc_scp3_accumulate.argtypes = [I, I, P, P, P, I, P, I, I, I, I64, I64, P, P, P, P, P]
        L.orc_grad_rule.argtypes = [I, P, P, P, P, D]
        L.orc_grad_rule.restype = D
        L.orc_grad_accumulate.argtypes = [I, I, I, I, I64, P, P, P, P]
        L.orc_qtip4pf_grad.argtypes = [I64, P, P]
        L.orc_qtip4pf_grad.restype = D
        L.orc_ttm3f_grad.argtypes = [I64, P, P, D, P]
        L.orc_ttm3f_grad.restype = D
        L.orc_scp3_accumulate_grad.argtypes = [I, I, P, P, P, I, P, I, I, I, I, I64, I64, P, P, P, P]
        L.orc_scp1_average.argtypes = [I, I, P, P, P, I64, I64, P, P, P]
        L.orc_block_weights.argtypes = [I, I, P, P, P, P, I64, P, P]
        L.orc_block_bra.argtypes = [I, P, I, I, I, I64, P, P]
        if hasattr(L, "orc_ttm3f"):
            try:
                L.orc_ttm3f.argtypes = [I64, P]
                L.orc_ttm3f.restype = D
            except AttributeError:
                pass
        _orc = L
    return _orc


def has_ref():
    return os.path.exists(REF)


def ref():
    """The reference's own code (oracle/_ref), or None if it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF):
        R = ctypes.CDLL(REF)
        R.ref_i8_sobol_block.argtypes = [I, ctypes.c_longlong, ctypes.c_longlong, P]
        R.ref_std_normal_block.argtypes = [I, ctypes.c_longlong, ctypes.c_longlong, P]
        R.ref_bsm.argtypes = [I, P, P]
        R.ref_ho.argtypes = [I, D]
        R.ref_ho.restype = D
        for nm in ("ref_qtip4pf", "ref_ttm3f"):
            f = getattr(R, nm)
            f.argtypes = [ctypes.c_size_t, P]
            f.restype = D
        for nm in ("ref_qtip4pf_grad", "ref_ttm3f_grad"):
            f = getattr(R, nm)
            f.argtypes = [ctypes.c_size_t, P, P]
            f.restype = D
        R.ref_pot_nasa.argtypes = [P]
        R.ref_pot_nasa.restype = D
        R.ref_dms_nasa.argtypes = [D, D, D, P, P, I]
        R.ref_smear012.argtypes = [D, D, P]
        _ref = R
    return _ref


# ---- numpy-level helpers ---------------------------------------------------

def sobol_ints(dims, seed0, npts):
    out = np.zeros((npts, dims), dtype=np.uint64)
    rc = lib().orc_sobol_ints(dims, seed0, npts, _p(out))
    assert rc == 0, rc
    return out


def sobol_normals(dims, seed0, npts, halve=True, scramble=None):
    """Normals of seeds seed0.. (points x dims); scramble = a seed for the
    affine digit scramble (None: the reference's sequence)."""
    y = np.zeros((npts, dims))
    if scramble is None:
        rc = lib().orc_sobol_normals(dims, seed0, npts, 1 if halve else 0, _p(y))
    else:
        rc = lib().orc_sobol_normals_scr(dims, seed0, npts, 1 if halve else 0, 1, int(scramble), _p(y))
    assert rc == 0, rc
    return y


def sobol_scramble(ints, seed):
    """Affine digit scramble of Sobol integers (points x dims), a copy."""
    x = np.ascontiguousarray(ints, dtype=np.uint64).copy()
    lib().orc_sobol_scramble(x.shape[1], int(seed), x.shape[0], _p(x))
    return x


def bsm(u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    x = np.zeros_like(u)
    lib().orc_bsm(u.size, _p(u), _p(x))
    return x


def basis_size(nsel, n2, n3):
    return lib().orc_basis_size(nsel, n2, n3)


def get_bra(q, nsel, n2, n3):
    """bra for each row of q (points x nsel) -> (points x n)."""
    q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
    n = basis_size(nsel, n2, n3)
    out = np.zeros((q.shape[0], n))
    L = lib()
    for i in range(q.shape[0]):
        row = np.ascontiguousarray(q[i])
        r = np.zeros(n)
        L.orc_get_bra(nsel, n2, n3, _p(row), _p(r))
        out[i] = r
    return out


def add_epot(bra, V, M):
    """M (n x n, Fortran order) upper += bra^T diag(V) bra for bra (P x n)."""
    bra = np.ascontiguousarray(bra, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    assert M.flags.f_contiguous
    lib().orc_add_epot(bra.shape[1], bra.shape[0], _p(bra), _p(V), _p(M))


def add_ho_diagonal(nsel, n2, n3, omega_sel, M):
    om = np.ascontiguousarray(omega_sel, dtype=np.float64)
    lib().orc_add_ho_diagonal(nsel, n2, n3, _p(om), M.shape[0], _p(M))


def qtip4pf(crd):
    c = np.ascontiguousarray(crd, dtype=np.float64).ravel()
    return lib().orc_qtip4pf(c.size // 9, _p(c))


def ttm3f(crd):
    c = np.ascontiguousarray(crd, dtype=np.float64).ravel()
    return lib().orc_ttm3f(c.size // 9, _p(c))


def scp3_accumulate(sysd, pes, seed0, npts, normals=None, want=("M",)):
    """Run the -G block loop (SCP3.cpp:187-232) on the CPU.

    sysd: dict from scp3_amd.system.prepare.  Returns dict with the raw sum M
    (n x n Fortran, upper), and y (points x M), V (hartree), bV on request.
    """
    L = lib()
    nw = sysd["nw"]
    M0 = sysd["omega"].size
    modes = np.ascontiguousarray(sysd["modes"], dtype=np.int32)
    n = basis_size(modes.size, sysd["nmodes2"], sysd["nmodes3"])
    out = {}
    Mm = np.zeros((n, n), order="F") if "M" in want else None
    y = np.zeros((npts, M0)) if "y" in want else None
    V = np.zeros(npts) if "V" in want else None
    bV = np.zeros(npts) if "bV" in want else None
    x0 = np.ascontiguousarray(sysd["x0"], dtype=np.float64)
    om = np.ascontiguousarray(sysd["omega"], dtype=np.float64)
    MU = np.asfortranarray(sysd["MUa"], dtype=np.float64)
    nr = np.ascontiguousarray(normals, dtype=np.float64) if normals is not None else None
    rc = L.orc_scp3_accumulate(nw, M0, _p(x0), _p(om), _p(MU), modes.size, _p(modes),
                               sysd["nmodes2"], sysd["nmodes3"], {"qtip4pf": 0, "ttm3f": 1}[pes],
                               seed0, npts, _p(nr), _p(Mm), _p(y), _p(V), _p(bV))
    assert rc == 0, rc
    out.update(M=Mm, y=y, V=V, bV=bV)
    return out


def qtip4pf_grad(crd):
    """q-TIP4P/F energy (kcal/mol) and gradient (kcal/mol/A) of 9nw coordinates."""
    c = np.ascontiguousarray(crd, dtype=np.float64).ravel()
    g = np.zeros_like(c)
    e = lib().orc_qtip4pf_grad(c.size // 9, _p(c), _p(g))
    return e, g


def ttm3f_grad(crd, tol=1e-15):
    """TTM3-F energy (kcal/mol) and gradient (kcal/mol/A) of 9nw coordinates."""
    c = np.ascontiguousarray(crd, dtype=np.float64).ravel()
    g = np.zeros_like(c)
    e = lib().orc_ttm3f_grad(c.size // 9, _p(c), _p(g), tol, None)
    return e, g


def grad_accumulate(q, V, vq, nsel, n2, n3, form=0):
    """Raw gradient-estimator sums (n x n Fortran, upper) for points q, V, vq.

    q, vq: points x nsel (selected modes); V: the weights bV.  form 0 is the
    reference's element placement, 1 the Stein rule for every element.
    """
    q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
    vq = np.ascontiguousarray(np.atleast_2d(vq), dtype=np.float64)
    V = np.ascontiguousarray(np.atleast_1d(V), dtype=np.float64)
    n = basis_size(nsel, n2, n3)
    M = np.zeros((n, n), order="F")
    rc = lib().orc_grad_accumulate(nsel, n2, n3, form, q.shape[0], _p(q), _p(V), _p(vq), _p(M))
    assert rc == 0, rc
    return M


def scp3_accumulate_grad(sysd, seed0, npts, form=0, normals=None, pes="qtip4pf"):
    """The gradient-estimator block loop (SCP3.cpp:187-232) with q-TIP4P/F or TTM3-F.

    Returns dict with the raw sums M (n x n Fortran, upper), bV and Vq
    (points x selected modes).
    """
    L = lib()
    nw = sysd["nw"]
    M0 = sysd["omega"].size
    modes = np.ascontiguousarray(sysd["modes"], dtype=np.int32)
    n = basis_size(modes.size, sysd["nmodes2"], sysd["nmodes3"])
    Mm = np.zeros((n, n), order="F")
    bV = np.zeros(npts)
    Vq = np.zeros((npts, modes.size))
    x0 = np.ascontiguousarray(sysd["x0"], dtype=np.float64)
    om = np.ascontiguousarray(sysd["omega"], dtype=np.float64)
    MU = np.asfortranarray(sysd["MUa"], dtype=np.float64)
    nr = np.ascontiguousarray(normals, dtype=np.float64) if normals is not None else None
    rc = L.orc_scp3_accumulate_grad(nw, M0, _p(x0), _p(om), _p(MU), modes.size, _p(modes),
                                    sysd["nmodes2"], sysd["nmodes3"], form, {"qtip4pf": 0, "ttm3f": 1}[pes],
                                    seed0, npts, _p(nr),
                                    _p(Mm), _p(bV), _p(Vq))
    assert rc == 0, rc
    return dict(M=Mm, bV=bV, Vq=Vq)


def scp1_average(pes, q, d, U, skip, npts):
    """SCP1::average (SCP1.cpp:33-86): <V>, <UX>, <UX q^T> in Hartree / bohr
    over the Sobol points skip .. skip + npts - 1 (scp1_oracle.c)."""
    q = np.ascontiguousarray(q, dtype=np.float64).ravel()
    d = np.ascontiguousarray(d, dtype=np.float64).ravel()
    Uf = np.asfortranarray(U, dtype=np.float64)
    n3 = q.size
    UX = np.zeros(n3)
    UXq = np.zeros((n3, n3), order="F")
    Up = ctypes.c_double()
    pid = {"qtip4pf": 0, "ttm3f": 1}[pes]
    rc = lib().orc_scp1_average(n3 // 9, pid, _p(q), _p(d), ctypes.c_void_p(Uf.ctypes.data), int(skip),
                                int(npts), ctypes.byref(Up), _p(UX), ctypes.c_void_p(UXq.ctypes.data))
    if rc:
        raise RuntimeError("orc_scp1_average: %d" % rc)
    return Up.value, UX, UXq


def finalize(sysd, Mraw, npts):
    """Mout = M/P + HO diagonal + unselected ZPE (SCP3.cpp:237-239,269-271)."""
    modes = sysd["modes"]
    om_sel = sysd["omega"][modes]
    M = Mraw / npts
    M = np.asfortranarray(M)
    add_ho_diagonal(modes.size, sysd["nmodes2"], sysd["nmodes3"], om_sel, M)
    M[np.diag_indices_from(M)] += 0.5 * (sysd["omega"].sum() - om_sel.sum())
    return M


def pes_pointer(name, reference=True):
    """Address of a C energy function with the PES.h signature (nw, crd) ->
    kcal/mol: the reference's own build (oracle/_ref) when present and
    `reference`, else this oracle's restatement.  Returns (address, kind)."""
    R = ref() if reference else None
    if R is not None:
        return ctypes.cast(getattr(R, "ref_" + name), ctypes.c_void_p).value, "reference"
    return ctypes.cast(getattr(lib(), "orc_" + name), ctypes.c_void_p).value, "port"


def block_weights(sysd, fnptr, y):
    """bV of a block of halved normals y (points x nmodes0), the serial
    potential loop of SCP3.cpp:205-225 through the C function fnptr."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    bV = np.zeros(y.shape[0])
    x0 = np.ascontiguousarray(sysd["x0"], dtype=np.float64)
    om = np.ascontiguousarray(sysd["omega"], dtype=np.float64)
    MU = np.asfortranarray(sysd["MUa"], dtype=np.float64)
    rc = lib().orc_block_weights(sysd["nw"], om.size, _p(x0), _p(om), _p(MU), ctypes.c_void_p(fnptr),
                                 y.shape[0], _p(y), _p(bV))
    assert rc == 0, rc
    return bV


def block_bra(sysd, y, out=None):
    """Basis rows of a block (points x n), get_bra under OpenMP."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    modes = np.ascontiguousarray(sysd["modes"], dtype=np.int32)
    n = basis_size(modes.size, sysd["nmodes2"], sysd["nmodes3"])
    if out is None:
        out = np.zeros((y.shape[0], n))
    rc = lib().orc_block_bra(modes.size, _p(modes), y.shape[1], sysd["nmodes2"], sysd["nmodes3"],
                             y.shape[0], _p(y), _p(out))
    assert rc == 0, rc
    return out
