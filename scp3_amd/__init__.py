"""MI355X Hamiltonian-assembly path of SCP3: Python bindings over libscp3g.

The product is the C ABI in include/scp3g.h (libscp3g.so, HIP kernels for
gfx950).  This module loads it with ctypes and wraps the context in a small
class; it has no compute path of its own, and raises if the library is
missing or no GPU is visible.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscp3g.so")

FP64 = 0
MIXED = 1
PES_QTIP4PF = 0
PES_TTM3F = 1
EST_ENERGY = 0          # -G: B^T diag(bV) B
EST_GRADIENT = 1        # the reference's default estimator (bV and its gradient)
EST_GRADIENT_STEIN = 2  # the same with every element from the Stein identity

# Constants.cpp:12-18, Constants.h:21
BOHR = 0.52917721092
AUTOCM = 2.194746313e5
AUTOKCALPMOL = 627.509474277194
HMASS = 1837.15137
CMASS = 21891.6543
OMASS = CMASS * 15.9994 / 12.0107

_lib = None

_SIGS = {
    "scp3g_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "scp3g_destroy": (None, [ctypes.c_void_p]),
    "scp3g_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "scp3g_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "scp3g_set_system": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "scp3g_set_basis": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int]),
    "scp3g_basis_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "scp3g_sub_basis_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
    "scp3g_set_pes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "scp3g_set_precision": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "scp3g_set_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "scp3g_set_mixed_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "scp3g_set_cv_terms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "scp3g_set_estimator": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "scp3g_set_scramble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64]),
    "scp3g_hessian": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_double, ctypes.c_void_p]),
    "scp3g_scp1_average": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "scp3g_pes_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_int64]),
    "scp3g_accumulate_surrogate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]),
    "scp3g_accumulate_correction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]),
    "scp3g_hold_as_offset": (ctypes.c_int, [ctypes.c_void_p]),
    "scp3g_clear_offset": (ctypes.c_int, [ctypes.c_void_p]),
    "scp3g_points_coords": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]),
    "scp3g_accumulate_energies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]),
    "scp3g_accumulate_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                                  ctypes.c_void_p]),
    "scp3g_reset": (ctypes.c_int, [ctypes.c_void_p]),
    "scp3g_accumulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]),
    "scp3g_accumulate_normals": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "scp3g_points_done": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "scp3g_h_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "scp3g_lowest_levels": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "scp3g_lowest_levels_sub": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_void_p]),
    "scp3g_matrix_lowest_levels": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int,
                                                  ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "scp3g_sobol_normals": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "scp3g_sobol_normals_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                                   ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "scp3g_map_coords": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]),
    "scp3g_pes_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                        ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]),
    "scp3g_basis_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]),
    "scp3g_h_accumulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p,
                                          ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    "scp3g_normal_modes": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "scp3g_pes_cpu": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]),
    "scp3g_mlp_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64]),
    "scp3g_mlp_param_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "scp3g_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "scp3g_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "scp3g_mlp_load_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "scp3g_mlp_sample_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]),
    "scp3g_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_void_p]),
    "scp3g_mlp_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "scp3g_mlp_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_void_p]),
    "scp3g_accumulate_cv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                           ctypes.c_int64]),
    "scp3g_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "scp3g_kernel_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_int64)]),
}


def exported_symbols():
    """Names the C ABI exports (include/scp3g.h)."""
    return sorted(_SIGS)


def lib():
    """Load libscp3g.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libscp3g.so is not built: run python -m scp3_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Scp3Error(RuntimeError):
    pass


class Context:
    """One libscp3g context (device, stream, system, basis, running H)."""

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.scp3g_create(ctypes.byref(h), device)
        if rc != 0:
            raise Scp3Error("scp3g_create failed (%d): no usable GPU %d" % (rc, device))
        self._h = h
        self.L = L

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            self.L.scp3g_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc):
        if rc != 0:
            msg = self.L.scp3g_last_error(self._h)
            raise Scp3Error("libscp3g error %d: %s" % (rc, msg.decode() if msg else ""))

    # set-up -------------------------------------------------------------
    def set_system(self, x0_bohr, omega, MUa):
        x0 = np.ascontiguousarray(x0_bohr, dtype=np.float64)
        om = np.ascontiguousarray(omega, dtype=np.float64)
        MU = np.asfortranarray(MUa, dtype=np.float64)
        nw = x0.size // 9
        self.check(self.L.scp3g_set_system(self._h, nw, _ptr(x0), om.size, _ptr(om), _ptr(MU)))
        self.nw, self.nmodes0 = nw, om.size

    def set_basis(self, modes, nmodes2, nmodes3):
        m = np.ascontiguousarray(modes, dtype=np.int32)
        self.check(self.L.scp3g_set_basis(self._h, m.size, _ptr(m), nmodes2, nmodes3))

    def basis_size(self):
        n = ctypes.c_int64()
        self.check(self.L.scp3g_basis_size(self._h, ctypes.byref(n)))
        return n.value

    def sub_basis_size(self, level):
        n = ctypes.c_int64()
        self.check(self.L.scp3g_sub_basis_size(self._h, level, ctypes.byref(n)))
        return n.value

    def set_pes(self, name):
        self.check(self.L.scp3g_set_pes(self._h, name.encode()))

    def set_precision(self, mode):
        self.check(self.L.scp3g_set_precision(self._h, mode))

    def set_tile(self, tile):
        self.check(self.L.scp3g_set_tile(self._h, tile))

    def set_estimator(self, est):
        """EST_ENERGY (-G), EST_GRADIENT or EST_GRADIENT_STEIN; clears the running sum."""
        self.check(self.L.scp3g_set_estimator(self._h, int(est)))

    def set_scramble(self, seed):
        """Affine digit scrambling of the Sobol points drawn from `seed` (any
        integer, 0 included); None restores the reference's sequence."""
        if seed is None:
            self.check(self.L.scp3g_set_scramble(self._h, 0, 0))
        else:
            self.check(self.L.scp3g_set_scramble(self._h, 1, int(seed)))

    def set_mixed_kernel(self, kernel):
        """0: 64 x 64 direct-load SYRKs; 1: LDS-tiled SYRKs (see scp3g.h)"""
        self.check(self.L.scp3g_set_mixed_kernel(self._h, int(kernel)))

    # whole path ---------------------------------------------------------
    def reset(self):
        self.check(self.L.scp3g_reset(self._h))

    def accumulate(self, skip, npts):
        self.check(self.L.scp3g_accumulate(self._h, int(skip), int(npts)))

    def accumulate_normals(self, normals):
        a = np.ascontiguousarray(normals, dtype=np.float64)
        self.check(self.L.scp3g_accumulate_normals(self._h, _ptr(a), a.shape[0]))

    def points_done(self):
        n = ctypes.c_int64()
        self.check(self.L.scp3g_points_done(self._h, ctypes.byref(n)))
        return n.value

    def h_download(self, raw=False):
        n = self.basis_size()
        M = np.zeros((n, n), dtype=np.float64, order="F")
        self.check(self.L.scp3g_h_download(self._h, _ptr(M), 1 if raw else 0))
        return M

    def lowest_levels(self, k, tol=0.0, vectors=False):
        n = self.basis_size()
        ev = np.zeros(k)
        V = np.zeros((n, k), order="F") if vectors else None
        self.check(self.L.scp3g_lowest_levels(self._h, k, tol, _ptr(ev), _ptr(V)))
        return (ev, V) if vectors else ev

    def lowest_levels_sub(self, nsub, k, tol=0.0):
        ev = np.zeros(k)
        self.check(self.L.scp3g_lowest_levels_sub(self._h, nsub, k, tol, _ptr(ev)))
        return ev

    def matrix_lowest_levels(self, M, k, tol=0.0, vectors=False):
        A = np.asfortranarray(M, dtype=np.float64)
        n = A.shape[0]
        ev = np.zeros(k)
        V = np.zeros((n, k), order="F") if vectors else None
        self.check(self.L.scp3g_matrix_lowest_levels(self._h, n, _ptr(A), k, tol, _ptr(ev), _ptr(V)))
        return (ev, V) if vectors else ev

    def sync(self):
        self.check(self.L.scp3g_sync(self._h))

    # surrogate / control variate -------------------------------------------
    def mlp_init(self, hidden, seed=0):
        self.check(self.L.scp3g_mlp_init(self._h, hidden, seed))

    def mlp_params(self):
        n = ctypes.c_int64()
        self.check(self.L.scp3g_mlp_param_count(self._h, ctypes.byref(n)))
        P = np.zeros(n.value, dtype=np.float32)
        nrm = np.zeros(3)
        self.check(self.L.scp3g_mlp_get_params(self._h, _ptr(P), _ptr(nrm)))
        return P, nrm

    def mlp_set_params(self, P, nrm=None):
        P = np.ascontiguousarray(P, dtype=np.float32)
        nr = np.ascontiguousarray(nrm, dtype=np.float64) if nrm is not None else None
        self.check(self.L.scp3g_mlp_set_params(self._h, _ptr(P), _ptr(nr)))

    def mlp_load_data(self, y, t):
        y = np.ascontiguousarray(y, dtype=np.float64)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self.check(self.L.scp3g_mlp_load_data(self._h, _ptr(y), _ptr(t), y.shape[0]))

    def mlp_sample_data(self, skip, npts):
        self.check(self.L.scp3g_mlp_sample_data(self._h, int(skip), int(npts)))

    def mlp_train(self, steps, batch, lr):
        loss = np.zeros(steps, dtype=np.float32)
        self.check(self.L.scp3g_mlp_train(self._h, steps, batch, lr, _ptr(loss)))
        return loss

    def mlp_grad(self, batch):
        n = ctypes.c_int64()
        self.check(self.L.scp3g_mlp_param_count(self._h, ctypes.byref(n)))
        g = np.zeros(n.value, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self.check(self.L.scp3g_mlp_grad(self._h, batch, _ptr(g), _ptr(loss)))
        return g, float(loss[0])

    def accumulate_surrogate(self, skip, npts):
        """surrogate term alone: sum B B^T w_nn over Sobol seeds skip.., normaliser += npts"""
        self.check(self.L.scp3g_accumulate_surrogate(self._h, int(skip), int(npts)))

    def accumulate_correction(self, skip, npts):
        """correction term alone: sum B B^T (bV - w_nn), normaliser += npts"""
        self.check(self.L.scp3g_accumulate_correction(self._h, int(skip), int(npts)))

    def hold_as_offset(self):
        """the running sum, normalised, becomes a fixed term of every later
        Mout and level solve; the running sum restarts at zero"""
        self.check(self.L.scp3g_hold_as_offset(self._h))

    def clear_offset(self):
        self.check(self.L.scp3g_clear_offset(self._h))

    def accumulate_cv(self, skip_nn, P_nn, skip_t, P_t):
        self.check(self.L.scp3g_accumulate_cv(self._h, int(skip_nn), int(P_nn), int(skip_t), int(P_t)))

    def points_coords(self, skip, npts):
        """Cartesians (A, npts x 9nw) of Sobol seeds skip.. for a CPU potential"""
        crd = np.zeros((npts, 9 * self.nw), dtype=np.float64)
        self.check(self.L.scp3g_points_coords(self._h, int(skip), int(npts), _ptr(crd)))
        return crd

    def accumulate_energies(self, skip, E):
        """accumulate Sobol seeds skip.. with caller energies E (kcal/mol)"""
        E = np.ascontiguousarray(E, dtype=np.float64)
        self.check(self.L.scp3g_accumulate_energies(self._h, int(skip), E.size, _ptr(E)))

    def accumulate_gradients(self, skip, E, G):
        """Caller's energies (kcal/mol) and gradients (points x 9nw, kcal/mol/A)
        for the points of points_coords(skip, len(E)), into a gradient estimator."""
        E = np.ascontiguousarray(E, dtype=np.float64)
        G = np.ascontiguousarray(G, dtype=np.float64)
        self.check(self.L.scp3g_accumulate_gradients(self._h, int(skip), E.size, _ptr(E), _ptr(G)))

    def scp1_average(self, pes, q, d, isqrtM_U, skip, npts):
        """SCP1's averages <V>, <UX>, <UX q^T> (Hartree, bohr) over the Sobol
        points skip .. skip + npts - 1; q, d (9nw), isqrtM_U (9nw x 9nw)."""
        q = np.ascontiguousarray(q, dtype=np.float64).ravel()
        d = np.ascontiguousarray(d, dtype=np.float64).ravel()
        n3 = q.size
        Um = np.asfortranarray(isqrtM_U, dtype=np.float64)
        if d.size != n3 or Um.shape != (n3, n3) or n3 % 9:
            raise ValueError("q, d: 9nw values; isqrtM_U: 9nw x 9nw")
        pid = {"qtip4pf": PES_QTIP4PF, "ttm3f": PES_TTM3F}[pes] if isinstance(pes, str) else int(pes)
        U = ctypes.c_double()
        UX = np.zeros(n3)
        UXq = np.zeros((n3, n3), order="F")
        self.check(self.L.scp3g_scp1_average(self._h, pid, n3 // 9, _ptr(q), _ptr(d), ctypes.c_void_p(Um.ctypes.data),
                                             int(skip), int(npts), ctypes.byref(U), _ptr(UX),
                                             ctypes.c_void_p(UXq.ctypes.data)))
        return U.value, UX, UXq

    def hessian(self, pes, q, s=1e-3):
        """Cartesian Hessian (Hartree/bohr^2) at q (bohr) by central
        differences of the gradient, step s bohr (getHessian)."""
        q = np.ascontiguousarray(q, dtype=np.float64).ravel()
        if q.size % 9:
            raise ValueError("q: 9nw values")
        pid = {"qtip4pf": PES_QTIP4PF, "ttm3f": PES_TTM3F}[pes] if isinstance(pes, str) else int(pes)
        H = np.zeros((q.size, q.size), order="F")
        self.check(self.L.scp3g_hessian(self._h, pid, q.size // 9, _ptr(q), float(s), ctypes.c_void_p(H.ctypes.data)))
        return H

    def accumulate_surrogate(self, skip, npts):
        self.check(self.L.scp3g_accumulate_surrogate(self._h, int(skip), int(npts)))

    def set_cv_terms(self, terms):
        self.check(self.L.scp3g_set_cv_terms(self._h, int(terms)))

    # timing -------------------------------------------------------------
    def set_timing(self, on):
        self.check(self.L.scp3g_set_timing(self._h, 1 if on else 0))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        self.check(self.L.scp3g_kernel_time(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value


def pes_cpu(name, crd, grad=False):
    """FP64 CPU energy (kcal/mol) of one configuration (Angstrom, O H H order)."""
    L = lib()
    x = np.ascontiguousarray(crd, dtype=np.float64).ravel()
    E = np.zeros(1)
    g = np.zeros_like(x) if grad else None
    rc = L.scp3g_pes_cpu(name.encode(), x.size // 9, _ptr(x), _ptr(E), _ptr(g))
    if rc != 0:
        raise Scp3Error("scp3g_pes_cpu(%s) failed (%d)" % (name, rc))
    return (E[0], g) if grad else E[0]


def normal_modes(mass, hess):
    """omega (hartree), MUa (3N x M col-major), omega_all = sqrt|ev| (SCP3.cpp:292-318)."""
    L = lib()
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    H = np.asfortranarray(hess, dtype=np.float64)
    n3 = mass.size
    om = np.zeros(n3 - 6)
    MU = np.zeros((n3, n3 - 6), order="F")
    oall = np.zeros(n3)
    rc = L.scp3g_normal_modes(n3, _ptr(mass), _ptr(H), _ptr(om), _ptr(MU), _ptr(oall))
    if rc != 0:
        raise Scp3Error("scp3g_normal_modes failed (%d)" % rc)
    return om, MU, oall
