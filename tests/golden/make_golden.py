"""Write the golden fixtures in tests/golden/ from the reference's own code.

The reference's CPU sources (sobol.cpp, beasley_springer_moro.cpp,
HarmonicOscillatorBasis.cpp, qtip4pf.cpp, ttm3f.cpp, ps.cpp,
ttm3f-bits-smear.cpp) are compiled from where they lie by oracle/Makefile
into oracle/_ref/libscp3ref.so; this script calls that build and stores
inputs and outputs as .npz data.  Run in a container that has the reference:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from scp3_amd import system  # noqa: E402


def _p(a):
    return oracle._p(a)


def sobol(R):
    out = {}
    # every dimension of the table over the first 2048 seeds (catch-up free)
    q = np.zeros((2048, 1111))
    R.ref_i8_sobol_block(1111, 0, 2048, _p(q))
    out["small_dims1111_seed0"] = (q * 2.0 ** 62).astype(np.uint64)
    # the reference's default skip 2^30 (SCP3.cpp:48), dims of the configs
    for dims in (12, 48, 174):
        y = np.zeros((64, dims))
        R.ref_std_normal_block(dims, 1 << 30, 64, _p(y))
        out["normals_skip30_dims%d" % dims] = y
        q = np.zeros((64, dims))
        R.ref_i8_sobol_block(dims, 1 << 30, 64, _p(q))
        out["ints_skip30_dims%d" % dims] = (q * 2.0 ** 62).astype(np.uint64)
        print("sobol dims", dims, flush=True)
    # an odd, unaligned start
    q = np.zeros((300, 48))
    R.ref_i8_sobol_block(48, 123457, 300, _p(q))
    out["ints_seed123457_dims48"] = (q * 2.0 ** 62).astype(np.uint64)
    return out


def bsm(R):
    u = np.concatenate([
        np.linspace(1e-12, 1 - 1e-12, 4001),
        np.array([0.08, 0.0799999, 0.08000001, 0.92, 0.9200001, 0.5, 2.0 ** -31, 1 - 2.0 ** -31]),
    ])
    x = np.zeros_like(u)
    uc = u.copy()  # the reference takes a non-const pointer; keep it alive
    R.ref_bsm(u.size, _p(uc), _p(x))
    ho = np.zeros((4, 201))
    ys = np.linspace(-4, 4, 201)
    for n in range(4):
        ho[n] = [R.ref_ho(n, float(v)) for v in ys]
    return {"bsm_u": u, "bsm_x": x, "ho_y": ys, "ho_vals": ho}


def potentials(R):
    out = {}
    rng = np.random.default_rng(1234)
    for name in ("dimer_qtip4pf", "hexamer_qtip4pf", "hexamer_ttm3f", "w20_ttm3f"):
        path = system.input_path(name)
        if not os.path.exists(path):
            continue
        species, x, _ = system.read_vladimir(path)
        _, xs, _, _ = system.ohhohh(species, x)
        cfg = xs[None, :] + 0.05 * rng.normal(size=(16, xs.size))
        cfg[0] = xs
        eq = np.zeros(16)
        et = np.zeros(16)
        for i in range(16):
            c = np.ascontiguousarray(cfg[i])
            eq[i] = R.ref_qtip4pf(c.size // 9, _p(c))
            et[i] = R.ref_ttm3f(c.size // 9, _p(c))
        out["crd_" + name] = cfg
        out["E_qtip4pf_" + name] = eq
        out["E_ttm3f_" + name] = et
        print("potentials", name, flush=True)
    return out


def main(parts):
    oracle.build(ref=True)
    R = oracle.ref()
    if R is None:
        sys.exit("the reference build oracle/_ref is missing")
    if "bsm" in parts:
        np.savez_compressed(os.path.join(HERE, "bsm_ho.npz"), **bsm(R))
    if "potentials" in parts:
        np.savez_compressed(os.path.join(HERE, "potentials.npz"), **potentials(R))
    if "sobol" in parts:
        np.savez_compressed(os.path.join(HERE, "sobol.npz"), **sobol(R))


if __name__ == "__main__":
    main(sys.argv[1:] or ["bsm", "potentials", "sobol"])
