// Host-side input/output shared by the SCP3 and eigensolver command-line
// drivers: the "vladimir" input format, the O H H permutation and the
// sM_*.h5 matrix files.
#pragma once

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#ifdef SCP3G_HAVE_HDF5
#include <hdf5.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <thread>
#endif

namespace cli {

// Constants.cpp:12-18
constexpr double bohr = 0.52917721092;
constexpr double autocm = 2.194746313e5;
constexpr double Hmass = 1837.15137;
constexpr double Dmass = 2.0 * Hmass;
constexpr double Cmass = 21891.6543;
constexpr double Omass = Cmass * 15.9994 / 12.0107;

// DiskIO.cpp:22-40
inline double nuclear_mass(const std::string &s)
{
    if (s == "H") return Hmass;
    if (s == "D") return Dmass;
    if (s == "C") return Cmass;
    if (s == "O") return Omass;
    return -1.0;
}

struct Input {
    std::vector<double> mass, x0, H;  // 3N, 3N (bohr), 3N x 3N column-major
    int n3 = 0;
};

// load_from_vladimir (DiskIO.cpp:72-116): atom count, one skipped line,
// species x y z (Angstrom), then H(j, i) for j <= i by columns.
inline bool load_vladimir(const std::string &name, Input &in, std::string &err)
{
    std::ifstream f(name.c_str());
    if (!f.good()) {
        err = "File not found: " + name;
        return false;
    }
    int N = 0;
    f >> N;
    std::string skip;
    std::getline(f, skip);
    std::getline(f, skip);
    in.n3 = 3 * N;
    in.mass.assign(in.n3, 0.0);
    in.x0.assign(in.n3, 0.0);
    for (int i = 0; i < N; ++i) {
        std::string sp;
        f >> sp;
        const double m = nuclear_mass(sp);
        for (int k = 0; k < 3; ++k) in.mass[3 * i + k] = m;
        f >> in.x0[3 * i] >> in.x0[3 * i + 1] >> in.x0[3 * i + 2];
    }
    for (double &v : in.x0) v /= bohr;
    in.H.assign((size_t)in.n3 * in.n3, 0.0);
    for (int i = 0; i < in.n3; ++i)
        for (int j = 0; j <= i; ++j) {
            double v;
            f >> v;
            if (!f.good() && !(f.eof() && i == in.n3 - 1 && j == i)) {
                in.H.clear();
                return true;  // the reference leaves H empty on a short read
            }
            in.H[(size_t)i * in.n3 + j] = in.H[(size_t)j * in.n3 + i] = v;
        }
    return true;
}

// OHHOHH (DiskIO.cpp:142-195): permute atoms into O H H blocks.
inline bool ohhohh(Input &in, std::string &err)
{
    const int n = in.n3;
    std::vector<int> p(n);
    int iO = 0, iH = 3;
    for (int i = 0; i < n;) {
        if (in.mass[i] == Omass) {
            for (int j = 0; j < 3; ++j) p[iO + j] = i + j;
            iO += 9;
            i += 3;
        } else if (in.mass[i] == Hmass || in.mass[i] == Dmass) {
            for (int j = 0; j < 3; ++j) p[iH + j] = i + j;
            iH += 3;
            iH += 3 * (iH % 9 == 0);
            i += 3;
        } else {
            err = "Uknown mass at i = " + std::to_string(i);
            return false;
        }
    }
    std::vector<double> m(n), x(n);
    for (int i = 0; i < n; ++i) {
        m[i] = in.mass[p[i]];
        x[i] = in.x0[p[i]];
    }
    in.mass.swap(m);
    in.x0.swap(x);
    if (!in.H.empty()) {
        std::vector<double> Hs((size_t)n * n);
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) Hs[(size_t)c * n + r] = in.H[(size_t)p[c] * n + p[r]];
        in.H.swap(Hs);
    }
    return true;
}

#ifdef SCP3G_HAVE_HDF5
// save_hdf5 (DiskIO.cpp:200-227): dataset "M", dims {n_cols, n_rows}, FP32 in
// the file, chunk {min(n_cols, ceil(131072 / n_rows)), n_rows}, deflate 6.
//
// The chunks are compressed by worker threads and stored with
// H5Dwrite_chunk: HDF5's deflate filter is zlib's compress2 at the
// dataset's level, so each stored chunk is the byte stream H5Dwrite would
// have produced in one thread (the edge chunk padded with zeros to the full
// chunk, as HDF5 stores it).  At 20825 states a file is 1.7 GB of FP32
// before compression, about 40 s of deflate in one thread.
inline bool save_hdf5(const std::string &name, const double *M, int64_t n, std::string &err)
{
    hid_t f = H5Fcreate(name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) {
        err = "cannot create " + name;
        return false;
    }
    constexpr int level = 6;
    hsize_t dims[2] = {(hsize_t)n, (hsize_t)n};
    hid_t sp = H5Screate_simple(2, dims, nullptr);
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    hsize_t chunk[2];
    chunk[0] = std::min<hsize_t>(dims[0], (128 * 1024 + dims[1] - 1) / dims[1]);
    chunk[1] = dims[1];
    H5Pset_chunk(pl, 2, chunk);
    H5Pset_deflate(pl, level);
    hid_t ds = H5Dcreate2(f, "M", H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
    bool ok = ds >= 0;
    if (ok && n > 0) {
        // dataset row i is column i of M (column-major: M[i n .. i n + n))
        const int64_t cr = (int64_t)chunk[0], nchunks = (n + cr - 1) / cr;
        const size_t raw = (size_t)cr * (size_t)n * sizeof(float);
        const int nthr = (int)std::max<unsigned>(1u, std::min<unsigned>(16u, std::thread::hardware_concurrency()));
        // batches of chunks bound the memory held for compressed output
        const int64_t batch = std::max<int64_t>(nthr, (int64_t)((size_t)1 << 30) / (int64_t)raw);
        std::vector<std::vector<unsigned char>> out((size_t)std::min(batch, nchunks));
        for (int64_t b0 = 0; ok && b0 < nchunks; b0 += batch) {
            const int64_t b1 = std::min(nchunks, b0 + batch);
            std::atomic<int64_t> next(b0);
            std::atomic<bool> zfail(false);
            auto work = [&]() {
                std::vector<float> buf((size_t)cr * n);
                for (int64_t k; (k = next.fetch_add(1)) < b1;) {
                    for (int64_t i = 0; i < cr; ++i) {
                        const int64_t c = k * cr + i;
                        float *dst = buf.data() + (size_t)i * n;
                        if (c < n) {
                            const double *src = M + (size_t)c * n;
                            for (int64_t r = 0; r < n; ++r) dst[r] = (float)src[r];
                        } else {
                            std::fill(dst, dst + n, 0.0f);
                        }
                    }
                    std::vector<unsigned char> &o = out[(size_t)(k - b0)];
                    uLongf zn = compressBound((uLong)raw);
                    o.resize(zn);
                    if (compress2(o.data(), &zn, reinterpret_cast<const Bytef *>(buf.data()), (uLong)raw, level) != Z_OK)
                        zfail = true;
                    o.resize(zn);
                }
            };
            std::vector<std::thread> th;
            const int nt = (int)std::min<int64_t>(nthr, b1 - b0);
            for (int t = 1; t < nt; ++t) th.emplace_back(work);
            work();
            for (auto &t : th) t.join();
            if (zfail) {
                ok = false;
                break;
            }
            for (int64_t k = b0; ok && k < b1; ++k) {
                const hsize_t off[2] = {(hsize_t)(k * cr), 0};
                const std::vector<unsigned char> &o = out[(size_t)(k - b0)];
                ok = H5Dwrite_chunk(ds, H5P_DEFAULT, 0, off, o.size(), o.data()) >= 0;
            }
        }
    }
    if (ds >= 0) H5Dclose(ds);
    H5Pclose(pl);
    H5Sclose(sp);
    H5Fclose(f);
    if (!ok) err = "cannot write " + name;
    return ok;
}

// load_hdf5 (DiskIO.cpp:267-289): M(dims[1], dims[0]) column-major, as double.
inline bool load_hdf5(const std::string &name, std::vector<double> &M, int64_t &n, std::string &err)
{
    hid_t f = H5Fopen(name.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) {
        err = "cannot open " + name;
        return false;
    }
    hid_t ds = H5Dopen2(f, "M", H5P_DEFAULT);
    if (ds < 0) {
        H5Fclose(f);
        err = "no dataset M in " + name;
        return false;
    }
    hid_t sp = H5Dget_space(ds);
    hsize_t dims[2] = {0, 0};
    const int nd = H5Sget_simple_extent_dims(sp, dims, nullptr);
    bool ok = nd == 2 && dims[0] == dims[1];
    if (!ok) err = "HDF5 dataset \"M\" has wrong shape";
    if (ok) {
        n = (int64_t)dims[0];
        M.assign((size_t)n * n, 0.0);
        ok = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, M.data()) >= 0;
        if (!ok) err = "cannot read " + name;
    }
    H5Sclose(sp);
    H5Dclose(ds);
    H5Fclose(f);
    return ok;
}
#endif

}  // namespace cli
