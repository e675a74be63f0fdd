"""Edge cases through the C ABI: empty and one-point inputs, a point count
one past a tile, the largest Sobol dimension count, and the error returns
of calls made out of order or with sizes out of range (each must fail with
a message and leave the context usable)."""
import ctypes

import numpy as np
import pytest

import scp3_amd
from oracle import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _vp(t):
    torch.cuda.synchronize()
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def ctx(gpu_ctx_factory):
    return gpu_ctx_factory(0)


def _setup(c, sysd, mode=scp3_amd.FP64, tile=1024):
    c.set_system(sysd["x0"], sysd["omega"], sysd["MUa"])
    c.set_scramble(None)
    c.set_basis(sysd["modes"], sysd["nmodes2"], sysd["nmodes3"])
    c.set_pes("qtip4pf")
    c.set_precision(mode)
    c.set_tile(tile)


@pytest.mark.parametrize("mode", [scp3_amd.FP64, scp3_amd.MIXED])
def test_empty_one_point_and_tile_plus_one(ctx, dimer, mode):
    """0 points leave the sum at zero and Mout undefined (an error); 1 point
    and tile + 1 points (a full tile and a one-point tile) equal the oracle."""
    print("edge case: precision", mode, flush=True)
    _setup(ctx, dimer, mode)
    ctx.reset()
    ctx.accumulate(1 << 30, 0)
    assert ctx.points_done() == 0
    assert not np.any(ctx.h_download(raw=True))
    with pytest.raises(scp3_amd.Scp3Error):
        ctx.h_download()
    tol = 1e-9 if mode == scp3_amd.FP64 else 2e-5
    for P in (1, 1025):
        ctx.reset()
        ctx.accumulate(12345, P)
        assert ctx.points_done() == P
        M = ctx.h_download(raw=True)
        ref = oracle.scp3_accumulate(dimer, "qtip4pf", 12345, P)["M"]
        assert np.abs(M - ref).max() <= tol * np.abs(ref).max(), (P, np.abs(M - ref).max())


def test_sobol_largest_dimension_count_and_limits(ctx):
    """1111 dimensions (the direction table's size) bit-exact against the
    oracle at an unaligned skip; 1112 dimensions, a negative count and a
    leading dimension shorter than the count are rejected."""
    dims, skip, P = 1111, 999, 700
    ld = 768
    y = torch.zeros(dims * ld, dtype=torch.float64, device="cuda")
    q = torch.zeros(dims * ld, dtype=torch.int64, device="cuda")
    ctx.check(ctx.L.scp3g_sobol_normals(ctx.handle, dims, skip, P, 1, _vp(y), ld, _vp(q)))
    ctx.sync()
    got = q.cpu().numpy().view(np.uint64).reshape(dims, ld)[:, :P].T
    np.testing.assert_array_equal(got, oracle.sobol_ints(dims, skip, P))
    for args in ((1112, 0, 10, ld), (12, 0, -1, ld), (12, 0, 800, ld)):
        for name in ("scp3g_sobol_normals", "scp3g_sobol_normals_noalloc"):
            rc = getattr(ctx.L, name)(ctx.handle, args[0], args[1], args[2], 1, _vp(y), args[3], None)
            assert rc != 0 and ctx.L.scp3g_last_error(ctx.handle)


def test_errors_leave_the_context_usable(ctx, dimer):
    """Out-of-order and out-of-range calls return an error with a message;
    the same context then runs the path normally."""
    with scp3_amd.Context(0) as c:
        with pytest.raises(scp3_amd.Scp3Error, match="system not set"):
            c.set_basis(np.arange(3, dtype=np.int32), 3, 0)
        with pytest.raises(scp3_amd.Scp3Error):
            c.accumulate(0, 10)
        c.set_system(dimer["x0"], dimer["omega"], dimer["MUa"])
        with pytest.raises(scp3_amd.Scp3Error, match="out of range"):
            c.set_basis(np.array([0, 12], dtype=np.int32), 2, 0)
        with pytest.raises(scp3_amd.Scp3Error, match="exceed"):
            c.set_basis(np.arange(4, dtype=np.int32), 5, 0)
        with pytest.raises(scp3_amd.Scp3Error):
            c.set_pes("whbb")
        with pytest.raises(scp3_amd.Scp3Error):
            c.set_tile(1000)
        with pytest.raises(scp3_amd.Scp3Error):
            c.accumulate_cv(0, 100, 0, 100)          # no surrogate initialised
        with pytest.raises(scp3_amd.Scp3Error):
            c.lowest_levels(29)
        _setup(c, dimer)
        c.reset()
        c.accumulate(1 << 30, 2000)
        ref = oracle.scp3_accumulate(dimer, "qtip4pf", 1 << 30, 2000)["M"]
        M = c.h_download(raw=True)
        assert np.abs(M - ref).max() <= 1e-9 * np.abs(ref).max()
