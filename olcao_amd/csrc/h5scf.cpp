// The setup's integral file in the reference's HDF5 layout, written through
// the HDF5 C library from the matrices the device keeps after
// olcao_setup_scf.  Layout restated from src/uolcao/hdf5SCFIntg.F90:105-460
// (groups, "%07d" names, packed dataspaces, chunking, deflate level 1,
// "status" attributes) and matrixSubs.F90:506-534 (column-packed upper
// triangle).  Host code only; the numbers come through the C-ABI.
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/olcao_h5.h"

namespace {

struct H5Error {
    std::string what;
};

hid_t check(hid_t id, const char *what)
{
    if (id < 0) throw H5Error{what};
    return id;
}

void check(herr_t e, const char *what, int)
{
    if (e < 0) throw H5Error{what};
}

// Closes an HDF5 handle when it leaves scope.
struct Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    ~Handle()
    {
        if (id >= 0) close(id);
    }
};

std::string name7(int i)
{
    char b[16];
    std::snprintf(b, sizeof b, "%07d", i);
    return b;
}

void set_status(hid_t group, int value = 1)
{
    const hsize_t one = 1;
    Handle space{check(H5Screate_simple(1, &one, nullptr), "attribute space"), H5Sclose};
    Handle attr{check(H5Acreate2(group, "status", H5T_NATIVE_INT, space.id, H5P_DEFAULT, H5P_DEFAULT),
                      "status attribute"),
                H5Aclose};
    check(H5Awrite(attr.id, H5T_NATIVE_INT, &value), "status write", 0);
}

// A chunked, deflated dataset of doubles (rank 1 or 2); dims in C order.
// data == nullptr creates it without writing (the reference creates every
// dataset of the file at set-up, also those of terms it does not compute,
// which then read as the fill value 0).  Returns the open dataset when keep
// is set (for an attribute), else -1.
hid_t write_dataset(hid_t group, const std::string &name, const hsize_t dims[2],
                    const hsize_t chunk[2], const double *data, int rank = 2, bool keep = false)
{
    Handle space{check(H5Screate_simple(rank, dims, nullptr), "dataspace"), H5Sclose};
    Handle dcpl{check(H5Pcreate(H5P_DATASET_CREATE), "dataset properties"), H5Pclose};
    check(H5Pset_chunk(dcpl.id, rank, chunk), "chunk", 0);
    check(H5Pset_deflate(dcpl.id, 1), "deflate", 0);
    Handle ds{check(H5Dcreate2(group, name.c_str(), H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, dcpl.id,
                               H5P_DEFAULT),
                    "dataset"),
              H5Dclose};
    if (data)
        check(H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "dataset write",
              0);
    if (!keep) return -1;
    const hid_t id = ds.id;
    ds.id = -1;
    return id;
}

}  // namespace

extern "C" olcao_status olcao_write_scf_integrals_h5(olcao_ctx *ctx, const char *path)
{
    if (!ctx || !path) return OLCAO_ERR_INPUT;
    long long dims[8];
    int flags[4];
    olcao_status st = olcao_get_dims(ctx, dims);
    if (st == OLCAO_OK) st = olcao_get_flags(ctx, flags);
    if (st != OLCAO_OK) return st;
    const int n = (int)dims[0], nc = (int)dims[1], nk = (int)dims[2], potDim = (int)dims[3];
    const bool gamma = flags[0] != 0;
    const int rel = flags[1];
    const int comps = gamma ? 1 : 2;                      // numComponents
    const size_t full = (size_t)comps * n * n * nk;
    const size_t npack = (size_t)n * (n + 1) / 2;
    std::vector<double> mat(full), packed(npack * comps);
    // a first read checks the setup has run before the file is touched
    if ((st = olcao_get_scf_term(ctx, 0, 0, mat.data())) != OLCAO_OK) return st;
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Handle file{check(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "file create"),
                    H5Fclose};
        Handle top{check(H5Gcreate2(file.id, "atomIntgGroup", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                         "atomIntgGroup"),
                   H5Gclose};
        // packed vale-vale dataspace (numComponents, n(n+1)/2) in Fortran
        // order; chunks of at most 250M elements (hdf5SCFIntg.F90:144-153)
        const hsize_t vvDims[2] = {npack, (hsize_t)comps};
        const hsize_t vvChunk[2] = {std::min<hsize_t>(npack, 250000000 / comps), (hsize_t)comps};
        auto write_term = [&](hid_t group, int which, int term) {
            if (olcao_get_scf_term(ctx, which, term, mat.data()) != OLCAO_OK)
                throw H5Error{"reading a term matrix from the device"};
            for (int k = 0; k < nk; ++k) {
                const double *m = mat.data() + (size_t)comps * n * n * k;
                size_t p = 0;
                for (int i = 0; i < n; ++i)            // column i, rows 0..i
                    for (int j = 0; j <= i; ++j, ++p)
                        for (int q = 0; q < comps; ++q)
                            packed[p * comps + q] = m[((size_t)j + (size_t)i * n) * comps + q];
                write_dataset(group, name7(k + 1), vvDims, vvChunk, packed.data());
            }
            set_status(group);
        };
        // the packed datasets of a term this path does not compute, and its
        // status 0 (the state initSCFIntegralHDF5 leaves them in,
        // hdf5SCFIntg.F90:380-530)
        auto empty_term = [&](hid_t group) {
            for (int k = 0; k < nk; ++k) write_dataset(group, name7(k + 1), vvDims, vvChunk, nullptr);
        };
        const struct {
            const char *name;
            int which;
        } terms[] = {{"atomOverlap", 0}, {"atomKEOverlap", 1}, {"atomMVOverlap", 2}, {"atomNPOverlap", 3}};
        for (const auto &t : terms) {
            Handle g{check(H5Gcreate2(top.id, t.name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), t.name),
                     H5Gclose};
            if (t.which != 2 || rel) {
                write_term(g.id, t.which, 0);
            } else {   // no mass-velocity term without rel
                empty_term(g.id);
                set_status(g.id, 0);
            }
        }
        // Operators outside this path, created empty with status 0 so that
        // the reference's reader (accessSCFIntegralHDF5, :563-760) finds
        // every group, dataset and attribute it opens: dipole and momentum
        // matrices (packed, one subgroup per Cartesian direction) and the
        // k-overlaps (full valeDim x valeDim, "real"/"imag" per k).
        hsize_t fullChunk[2] = {(hsize_t)n, (hsize_t)n};
        if ((double)n * n > 250000000.0) {
            const double f = std::sqrt((double)n * n / 250000000.0);
            fullChunk[0] = fullChunk[1] = std::max<hsize_t>(1, (hsize_t)(n / f));
        }
        const hsize_t fullDims[2] = {(hsize_t)n, (hsize_t)n};
        for (const char *name : {"atomDMOverlap", "atomMMOverlap", "atomKOverlap", "atomKOverlapPlusG"}) {
            Handle g{check(H5Gcreate2(top.id, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), name),
                     H5Gclose};
            const bool kov = name[4] == 'K';
            for (int x = 1; x <= 3; ++x) {
                Handle sub{check(H5Gcreate2(g.id, name7(x).c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                                 "xyz group"),
                           H5Gclose};
                for (int k = 0; k < nk; ++k) {
                    if (kov) {
                        write_dataset(sub.id, "real" + name7(k + 1), fullDims, fullChunk, nullptr);
                        write_dataset(sub.id, "imag" + name7(k + 1), fullDims, fullChunk, nullptr);
                    } else {
                        write_dataset(sub.id, name7(k + 1), vvDims, vvChunk, nullptr);
                    }
                }
            }
            set_status(g.id, 0);
        }
        {
            Handle pot{check(H5Gcreate2(top.id, "atomPotOverlap", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                             "atomPotOverlap"),
                       H5Gclose};
            for (int j = 0; j < potDim; ++j) {
                Handle g{check(H5Gcreate2(pot.id, name7(j + 1).c_str(), H5P_DEFAULT, H5P_DEFAULT,
                                          H5P_DEFAULT),
                               "potential term group"),
                         H5Gclose};
                write_term(g.id, 4, j);
            }
        }
        if (nc > 0) {
            // coreVale (coreDim, valeDim) per component, "real<k>" and
            // "imag<k>" (hdf5SCFIntg.F90:438-446); chunks as :165-175
            Handle g{check(H5Gcreate2(top.id, "atomOverlapCV", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                           "atomOverlapCV"),
                     H5Gclose};
            std::vector<double> cv((size_t)comps * nc * n * nk), part((size_t)nc * n);
            if (olcao_get_scf_term(ctx, 5, 0, cv.data()) != OLCAO_OK)
                throw H5Error{"reading coreVale from the device"};
            const hsize_t cvDims[2] = {(hsize_t)n, (hsize_t)nc};
            hsize_t cvChunk[2] = {(hsize_t)n, (hsize_t)nc};
            const double prod = (double)nc * n;
            if (prod > 250000000.0) {
                const double f = std::sqrt(prod / 250000000.0);
                cvChunk[0] = std::max<hsize_t>(1, (hsize_t)(n / f));
                cvChunk[1] = std::max<hsize_t>(1, (hsize_t)(nc / f));
            }
            for (int k = 0; k < nk; ++k)
                for (int q = 0; q < comps; ++q) {
                    const double *src = cv.data() + (size_t)comps * nc * n * k;
                    for (size_t e = 0; e < (size_t)nc * n; ++e) part[e] = src[e * comps + q];
                    write_dataset(g.id, (q == 0 ? "real" : "imag") + name7(k + 1), cvDims, cvChunk,
                                  part.data());
                }
        }
    } catch (const H5Error &e) {
        std::fprintf(stderr, "olcao_write_scf_integrals_h5: %s failed for %s\n", e.what.c_str(), path);
        return OLCAO_ERR_INPUT;
    }
    return OLCAO_OK;
}

extern "C" olcao_status olcao_write_scf_vector_terms_h5(olcao_ctx *ctx, int which,
                                                        const double *dipoleCenter,
                                                        const char *path)
{
    if (!ctx || !path || (which & ~3) || ((which & 1) && !dipoleCenter)) return OLCAO_ERR_INPUT;
    long long dims[8];
    int flags[4];
    olcao_status st = olcao_get_dims(ctx, dims);
    if (st == OLCAO_OK) st = olcao_get_flags(ctx, flags);
    if (st != OLCAO_OK) return st;
    const int n = (int)dims[0], nc = (int)dims[1], nk = (int)dims[2];
    const int comps = flags[0] ? 1 : 2;
    const size_t full = (size_t)comps * n * n * nk;
    const size_t npack = (size_t)n * (n + 1) / 2;
    std::vector<double> mat(3 * full), packed(npack * comps);
    std::vector<double> cv(std::max<size_t>(3 * (size_t)comps * nc * n * nk, 1)),
        cc(std::max<size_t>(3 * (size_t)comps * nc * nc * nk, 1));
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        if (H5Fis_hdf5(path) <= 0) throw H5Error{"opening a file that is not HDF5"};
        Handle file{check(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), "file open"), H5Fclose};
        for (int op = 0; op < 2; ++op) {
            if (!(which >> op & 1)) continue;
            const olcao_status r = op == 0
                ? olcao_gauss_overlap_dm(ctx, dipoleCenter, mat.data(), cv.data(), cc.data())
                : olcao_gauss_overlap_mm(ctx, mat.data(), cv.data(), cc.data());
            if (r != OLCAO_OK) return r;
            const std::string grp = std::string("/atomIntgGroup/") + (op == 0 ? "atomDMOverlap" : "atomMMOverlap");
            Handle g{check(H5Gopen2(file.id, grp.c_str(), H5P_DEFAULT), "term group"), H5Gclose};
            for (int x = 0; x < 3; ++x)
                for (int k = 0; k < nk; ++k) {
                    const double *m = mat.data() + x * full + (size_t)comps * n * n * k;
                    size_t p = 0;
                    for (int i = 0; i < n; ++i)            // column i, rows 0..i
                        for (int j = 0; j <= i; ++j, ++p)
                            for (int q = 0; q < comps; ++q)
                                packed[p * comps + q] = m[((size_t)j + (size_t)i * n) * comps + q];
                    const std::string ds = grp + "/" + name7(x + 1) + "/" + name7(k + 1);
                    Handle d{check(H5Dopen2(file.id, ds.c_str(), H5P_DEFAULT), "term dataset"), H5Dclose};
                    check(H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, packed.data()),
                          "term write", 0);
                }
            Handle a{check(H5Aopen(g.id, "status", H5P_DEFAULT), "status attribute"), H5Aclose};
            const int one = 1;
            check(H5Awrite(a.id, H5T_NATIVE_INT, &one), "status write", 0);
        }
    } catch (const H5Error &e) {
        std::fprintf(stderr, "olcao_write_scf_vector_terms_h5: %s failed for %s\n", e.what.c_str(), path);
        return OLCAO_ERR_INPUT;
    }
    return OLCAO_OK;
}

extern "C" olcao_status olcao_write_scf_eigen_h5(olcao_ctx *ctx, int numStates, const char *path)
{
    if (!ctx || !path) return OLCAO_ERR_INPUT;
    long long dims[8];
    int flags[4];
    olcao_status st = olcao_get_dims(ctx, dims);
    if (st == OLCAO_OK) st = olcao_get_flags(ctx, flags);
    if (st != OLCAO_OK) return st;
    const int n = (int)dims[0], nk = (int)dims[2];
    const bool gamma = flags[0] != 0;
    const int spin = flags[2], comps = gamma ? 1 : 2;
    std::vector<std::vector<double>> w(spin), v(spin);
    for (int q = 0; q < spin; ++q) {
        w[q].resize((size_t)nk * numStates);
        v[q].resize((size_t)nk * numStates * n * comps);
        if ((st = olcao_get_eigen(ctx, q, numStates, w[q].data(), v[q].data())) != OLCAO_OK) return st;
    }
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        // an existing HDF5 file gets the groups added; a path that exists
        // but is not one is left alone
        std::FILE *probe = std::fopen(path, "rb");
        hid_t fid;
        if (probe) {
            std::fclose(probe);
            if (H5Fis_hdf5(path) <= 0) throw H5Error{"opening a file that is not HDF5"};
            fid = check(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), "file open");
        } else {
            fid = check(H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), "file create");
        }
        Handle file{fid, H5Fclose};
        for (const char *g : {"eigenValues", "eigenVectors"})
            if (H5Lexists(file.id, g, H5P_DEFAULT) > 0) check(H5Ldelete(file.id, g, H5P_DEFAULT), "old group", 0);
        Handle gv{check(H5Gcreate2(file.id, "eigenValues", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                        "eigenValues"),
                  H5Gclose};
        Handle ge{check(H5Gcreate2(file.id, "eigenVectors", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                        "eigenVectors"),
                  H5Gclose};
        // (valeDim, numStates) in Fortran order; chunks as
        // hdf5SCFEigVec.F90:82-89
        const hsize_t vDims[2] = {(hsize_t)numStates, (hsize_t)n};
        hsize_t vChunk[2] = {(hsize_t)numStates, (hsize_t)n};
        if ((double)n * numStates > 250000000.0) {
            vChunk[0] = std::max<hsize_t>(1, (hsize_t)(std::sqrt(250000000.0 / ((double)n * numStates)) * numStates));
            vChunk[1] = std::max<hsize_t>(1, std::min<hsize_t>(n, 250000000 / vChunk[0]));
        }
        const hsize_t sDims[2] = {(hsize_t)numStates, 1}, sChunk[2] = {(hsize_t)numStates, 1};
        std::vector<double> part((size_t)numStates * n);
        for (int q = 0; q < spin; ++q)
            for (int k = 0; k < nk; ++k) {
                const std::string kq = name7(k + 1) + name7(q + 1);
                write_dataset(gv.id, kq, sDims, sChunk, w[q].data() + (size_t)k * numStates, 1);
                const double *src = v[q].data() + (size_t)k * numStates * n * comps;
                for (int c = 0; c < comps; ++c) {
                    for (size_t e = 0; e < part.size(); ++e) part[e] = src[e * comps + c];
                    const std::string name = gamma ? kq : (c == 0 ? "real" : "imag") + kq;
                    const hid_t ds = write_dataset(ge.id, name, vDims, vChunk, part.data(), 2, c == 0);
                    if (c == 0) {
                        Handle keep{ds, H5Dclose};
                        set_status(keep.id, 0);
                    }
                }
            }
    } catch (const H5Error &e) {
        std::fprintf(stderr, "olcao_write_scf_eigen_h5: %s failed for %s\n", e.what.c_str(), path);
        return OLCAO_ERR_INPUT;
    }
    return OLCAO_OK;
}
