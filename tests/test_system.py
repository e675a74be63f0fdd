"""Input handling: vladimir format, O H H permutation, normal modes."""
import numpy as np
import pytest

import scp3_amd
from scp3_amd import system


def test_vladimir_roundtrip(tmp_path):
    sp, x, H = system.read_vladimir(system.input_path("dimer_qtip4pf"))
    p = tmp_path / "d.dat"
    system.write_vladimir(str(p), sp, x, H)
    sp2, x2, H2 = system.read_vladimir(str(p))
    assert sp2 == sp
    np.testing.assert_array_equal(x2, x)
    np.testing.assert_array_equal(H2, H)


def test_ohhohh_permutation():
    # H O H ordering with a deuterium: O first, then the two hydrogens
    sp = ["H", "O", "D", "O", "H", "H"]
    x = np.arange(18, dtype=float)
    mass, xs, _, p = system.ohhohh(sp, x)
    assert list(xs[:3]) == [3, 4, 5] and list(xs[3:6]) == [0, 1, 2] and list(xs[6:9]) == [6, 7, 8]
    assert list(xs[9:12]) == [9, 10, 11]
    assert mass[0] == scp3_amd.OMASS and mass[6] == 2 * scp3_amd.HMASS
    with pytest.raises(ValueError):
        system.ohhohh(["X", "O", "H"], np.zeros(9))


@pytest.mark.parametrize("name", ["dimer_qtip4pf", "hexamer_qtip4pf"])
def test_normal_modes_match_lapack(name):
    sp, x, H = system.read_vladimir(system.input_path(name))
    mass, xs, Hs, _ = system.ohhohh(sp, x, H)
    om, MU, oall = scp3_amd.normal_modes(mass, Hs)
    ev, U = np.linalg.eigh(Hs)
    np.testing.assert_allclose(om, np.sqrt(ev[6:]), rtol=1e-9)
    np.testing.assert_allclose(oall[6:], np.sqrt(ev[6:]), rtol=1e-9)
    assert np.all(oall[:6] < 1e-4)  # translations / rotations projected out
    # MUa columns = U[:,6+k]/(sqrt(m) sqrt(omega_k)) up to sign
    ref = U[:, 6:] / (np.sqrt(mass)[:, None] * np.sqrt(np.sqrt(ev[6:]))[None, :])
    for k in range(om.size):
        s = np.sign(ref[:, k].dot(MU[:, k]))
        np.testing.assert_allclose(MU[:, k], s * ref[:, k], atol=1e-8 * np.abs(ref[:, k]).max())


def test_prepare_sizes(dimer, hexamer):
    assert dimer["nw"] == 2 and dimer["omega"].size == 12
    assert hexamer["nw"] == 6 and hexamer["omega"].size == 48
    # all frequencies real and ascending
    assert np.all(np.diff(hexamer["omega"]) >= 0) and hexamer["omega"][0] > 0


def test_inputs_are_minima():
    for name in ("dimer_qtip4pf", "hexamer_qtip4pf"):
        sp, x, H = system.read_vladimir(system.input_path(name))
        _, xs, _, _ = system.ohhohh(sp, x)
        E, g = scp3_amd.pes_cpu("qtip4pf", xs, grad=True)
        assert np.abs(g).max() < 1e-4


def test_ttm3f_inputs_are_minima():
    """The TTM3-F inputs (hexamer, 20-water cluster) sit at a minimum of
    TTM3-F: the oracle's analytic gradient vanishes there (kcal/mol/A) and
    every kept frequency of the input Hessian is real and positive."""
    from oracle import oracle
    for name, m in (("hexamer_ttm3f", 48), ("w20_ttm3f", 174)):
        sp, x, H = system.read_vladimir(system.input_path(name))
        _, xs, _, _ = system.ohhohh(sp, x)
        E, g = oracle.ttm3f_grad(xs)
        assert np.abs(g).max() < 1e-4, (name, np.abs(g).max())
        sysd = system.prepare(system.input_path(name), nmodes2=1)
        assert sysd["omega"].size == m and sysd["omega"][0] > 10.0 / scp3_amd.AUTOCM


def test_normals_file_round_trip(tmp_path):
    from scp3_amd import system
    rng = np.random.default_rng(4)
    Z = rng.normal(size=(37, 12))
    path = str(tmp_path / "normals.txt")
    system.write_normals(path, Z)
    np.testing.assert_array_equal(system.read_normals(path, 12), Z)
