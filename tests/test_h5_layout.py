"""The sM_*.h5 layout the drivers write, checked independently of their
reader: h5dump's view of the file against save_hdf5 (DiskIO.cpp:200-227).
Dataset "M", rank 2, dims {n, n}, file type IEEE FP32 little-endian, chunk
{min(n, ceil(131072 / n)), n}, deflate level 6, and the column-major FP64
Mout written as-is: dataset row c holds column c of M, so the upper-triangle
element M(r, c), r <= c, sits at [c][r] and the rest is zero."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDF5_INC = "/opt/conda/include"
HDF5_LIB_DIR = os.path.join(ROOT, "scp3_amd", "bin", "hdf5")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    if not os.path.exists(os.path.join(HDF5_INC, "hdf5.h")) or not os.path.isdir(HDF5_LIB_DIR):
        pytest.skip("HDF5 headers or the drivers' HDF5 library link are absent")
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump is absent")
    libs = [f for f in os.listdir(HDF5_LIB_DIR) if f.startswith("libhdf5.so")]
    exe = str(tmp_path_factory.mktemp("h5") / "h5_layout")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-DSCP3G_HAVE_HDF5", "-I" + HDF5_INC,
                        os.path.join(ROOT, "tests", "h5_layout_main.cpp"), "-o", exe,
                        os.path.join(HDF5_LIB_DIR, libs[0]), "-lz", "-Wl,-rpath," + HDF5_LIB_DIR],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("n", [91, 300, 1225, 1500])
def test_sm_h5_layout(writer, tmp_path, n):
    path = str(tmp_path / "sM_0000001.h5")
    r = subprocess.run([writer, str(n), path], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    hdr = subprocess.run([H5DUMP, "-p", "-H", path], capture_output=True, text=True)
    assert hdr.returncode == 0, hdr.stderr
    h = hdr.stdout
    assert 'DATASET "M"' in h
    assert "H5T_IEEE_F32LE" in h
    assert re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(\s*%d,\s*%d\s*\)" % (n, n), h), h
    c0 = min(n, -(-131072 // n))
    assert re.search(r"CHUNKED\s+\(\s*%d,\s*%d\s*\)" % (c0, n), h), h
    assert re.search(r"DEFLATE\s*\{\s*LEVEL\s+6\s*\}", h), h
    # values in dataset order: row c of the dataset = column c of M
    dat = subprocess.run([H5DUMP, "-d", "M", "-y", "-w", "0", "-m", "%.9g", path], capture_output=True, text=True)
    assert dat.returncode == 0, dat.stderr
    body = dat.stdout.split("DATA {", 1)[1].rsplit("}", 3)[0]
    vals = np.array([float(v) for v in re.split(r"[,\s]+", body.strip()) if v])
    assert vals.size == n * n
    D = vals.reshape(n, n)                     # D[c][r]
    c, r = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    expect = np.where(r <= c, r + 0.25 * c, 0.0)
    np.testing.assert_array_equal(D, expect)
