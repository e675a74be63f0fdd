"""scp3g_sobol_normals_noalloc (the entry point that allocates nothing)
against scp3g_sobol_normals on the same seeds: bit-identical normals and
Sobol integers on two calls in a row, the second with fewer dimensions and
points, halved and not; the integers also against the oracle."""
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


def test_sobol_normals_noalloc_matches_allocating_entry():
    with scp3_amd.Context(0) as ctx:
        for dims, skip, npts, halve in ((174, 1 << 30, 5000, 1), (48, 123457, 2049, 0), (12, 7, 300, 1)):
            print("no-alloc sobol case dims %d skip %d npts %d halve %d" % (dims, skip, npts, halve), flush=True)
            ld = (npts + 255) // 256 * 256
            out = {}
            for name in ("scp3g_sobol_normals", "scp3g_sobol_normals_noalloc"):
                y = torch.full((dims * ld,), 7.0, dtype=torch.float64, device="cuda")
                q = torch.zeros(dims * ld, dtype=torch.int64, device="cuda")
                ctx.check(getattr(ctx.L, name)(ctx.handle, dims, skip, npts, halve, _vp(y), ld, _vp(q)))
                ctx.sync()
                out[name] = (y.cpu().numpy().reshape(dims, ld), q.cpu().numpy().view(np.uint64).reshape(dims, ld))
            (y1, q1), (y2, q2) = out["scp3g_sobol_normals"], out["scp3g_sobol_normals_noalloc"]
            np.testing.assert_array_equal(q2[:, :npts], q1[:, :npts])
            np.testing.assert_array_equal(y2[:, :npts].view(np.uint64), y1[:, :npts].view(np.uint64))
            assert np.all(y2[:, npts:] == 7.0)
            np.testing.assert_array_equal(q2[:, :npts].T, oracle.sobol_ints(dims, skip, npts))
        with pytest.raises(scp3_amd.Scp3Error):
            y = torch.zeros(1112 * 256, dtype=torch.float64, device="cuda")
            ctx.check(ctx.L.scp3g_sobol_normals_noalloc(ctx.handle, 1112, 0, 256, 1, _vp(y), 256, None))
