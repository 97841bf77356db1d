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

// Column k of a [nk][n][n] (complex: interleaved) matrix set packed as
// matrixSubs.F90:506-534 does: column i, rows 0..i, components innermost.
void pack_upper(const double *mat, int n, int comps, int k, double *packed)
{
    const double *m = mat + (size_t)comps * n * n * k;
    size_t p = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j, ++p)
            for (int q = 0; q < comps; ++q)
                packed[p * comps + q] = m[((size_t)j + (size_t)i * n) * comps + q];
}

// The packed vale-vale dataspace (numComponents, n(n+1)/2) in Fortran order
// and its chunks of at most 250M elements (hdf5SCFIntg.F90:144-153,
// hdf5PSCFIntg.F90:123-133).
struct PackedSpace {
    hsize_t dims[2], chunk[2];
    PackedSpace(int n, int comps)
    {
        const hsize_t npack = (hsize_t)n * (n + 1) / 2;
        dims[0] = npack;
        dims[1] = (hsize_t)comps;
        chunk[0] = std::min<hsize_t>(npack, 250000000 / comps);
        chunk[1] = (hsize_t)comps;
    }
};

// One packed dataset per k-point of a term group, then its status 1.
void write_packed_term(hid_t group, const double *mat, int n, int comps, int nk)
{
    const PackedSpace sp(n, comps);
    std::vector<double> packed((size_t)n * (n + 1) / 2 * comps);
    for (int k = 0; k < nk; ++k) {
        pack_upper(mat, n, comps, k, packed.data());
        write_dataset(group, name7(k + 1), sp.dims, sp.chunk, packed.data());
    }
    set_status(group);
}

// The groups of operators outside this path, created with their datasets
// and status 0 so that the reference's readers find every group, dataset
// and attribute they open: dipole and momentum matrices (packed, one
// subgroup per Cartesian direction) and the k-overlaps (full valeDim x
// valeDim, "real"/"imag" per k).
void empty_operator_groups(hid_t top, int n, int comps, int nk)
{
    const PackedSpace sp(n, comps);
    hsize_t fullChunk[2] = {(hsize_t)n, (hsize_t)n};
    if ((double)n * n > 250000000.0) {
        const double f = std::sqrt((double)n * n / 250000000.0);
        fullChunk[0] = fullChunk[1] = std::max<hsize_t>(1, (hsize_t)(n / f));
    }
    const hsize_t fullDims[2] = {(hsize_t)n, (hsize_t)n};
    for (const char *name : {"atomDMOverlap", "atomMMOverlap", "atomKOverlap", "atomKOverlapPlusG"}) {
        Handle g{check(H5Gcreate2(top, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), name), H5Gclose};
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
                    write_dataset(sub.id, name7(k + 1), sp.dims, sp.chunk, nullptr);
                }
            }
        }
        set_status(g.id, 0);
    }
}

// coreVale (coreDim, valeDim) per component, "real<k>" and "imag<k>"
// (hdf5SCFIntg.F90:438-446), chunks as :165-175; cv: coreVale(coreDim,
// valeDim, nk) (complex interleaved) as olcao_gauss_overlap returns it.
void write_core_vale(hid_t top, const double *cv, int n, int nc, int comps, int nk)
{
    Handle g{check(H5Gcreate2(top, "atomOverlapCV", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                   "atomOverlapCV"),
             H5Gclose};
    std::vector<double> part((size_t)nc * n);
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
            const double *src = cv + (size_t)comps * nc * n * k;
            for (size_t e = 0; e < (size_t)nc * n; ++e) part[e] = src[e * comps + q];
            write_dataset(g.id, (q == 0 ? "real" : "imag") + name7(k + 1), cvDims, cvChunk,
                          part.data());
        }
}

// eigenValues/<k><s> and eigenVectors/{real,imag}<k><s> (one real dataset
// at Gamma) under `parent`, each first eigenvector dataset with status 0
// (hdf5SCFEigVal.f90:60-95, hdf5SCFEigVec.F90:70-160, and the same layout
// in hdf5PSCFEigVal.f90 and hdf5PSCFEigVec.F90).  w[q]: [nk][numStates];
// v[q]: [nk][numStates][n] (complex interleaved).
void write_eigen_groups(hid_t parent, const std::vector<std::vector<double>> &w,
                        const std::vector<std::vector<double>> &v, int n, int nk, int numStates,
                        bool gamma)
{
    const int comps = gamma ? 1 : 2;
    for (const char *g : {"eigenValues", "eigenVectors"})
        if (H5Lexists(parent, g, H5P_DEFAULT) > 0) check(H5Ldelete(parent, g, H5P_DEFAULT), "old group", 0);
    Handle gv{check(H5Gcreate2(parent, "eigenValues", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                    "eigenValues"),
              H5Gclose};
    Handle ge{check(H5Gcreate2(parent, "eigenVectors", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                    "eigenVectors"),
              H5Gclose};
    // (valeDim, numStates) in Fortran order; chunks as hdf5SCFEigVec.F90:82-89
    const hsize_t vDims[2] = {(hsize_t)numStates, (hsize_t)n};
    hsize_t vChunk[2] = {(hsize_t)numStates, (hsize_t)n};
    if ((double)n * numStates > 250000000.0) {
        vChunk[0] = std::max<hsize_t>(1, (hsize_t)(std::sqrt(250000000.0 / ((double)n * numStates)) * numStates));
        vChunk[1] = std::max<hsize_t>(1, std::min<hsize_t>(n, 250000000 / vChunk[0]));
    }
    const hsize_t sDims[2] = {(hsize_t)numStates, 1}, sChunk[2] = {(hsize_t)numStates, 1};
    std::vector<double> part((size_t)numStates * n);
    for (size_t q = 0; q < w.size(); ++q)
        for (int k = 0; k < nk; ++k) {
            const std::string kq = name7(k + 1) + name7((int)q + 1);
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
}

// An existing HDF5 file opened for writing, or a new one; a path that
// exists but holds another kind of file is left alone.
hid_t open_or_create(const char *path)
{
    std::FILE *probe = std::fopen(path, "rb");
    if (probe) {
        std::fclose(probe);
        if (H5Fis_hdf5(path) <= 0) throw H5Error{"opening a file that is not HDF5"};
        return check(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), "file open");
    }
    return check(H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), "file create");
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
    std::vector<double> mat(full);
    // a first read checks the setup has run before the file is touched
    if ((st = olcao_get_scf_term(ctx, 0, 0, mat.data())) != OLCAO_OK) return st;
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Handle file{check(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "file create"),
                    H5Fclose};
        Handle top{check(H5Gcreate2(file.id, "atomIntgGroup", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                         "atomIntgGroup"),
                   H5Gclose};
        const PackedSpace sp(n, comps);
        auto write_term = [&](hid_t group, int which, int term) {
            if (olcao_get_scf_term(ctx, which, term, mat.data()) != OLCAO_OK)
                throw H5Error{"reading a term matrix from the device"};
            write_packed_term(group, mat.data(), n, comps, nk);
        };
        // the packed datasets of a term this path does not compute, and its
        // status 0 (the state initSCFIntegralHDF5 leaves them in,
        // hdf5SCFIntg.F90:380-530)
        auto empty_term = [&](hid_t group) {
            for (int k = 0; k < nk; ++k) write_dataset(group, name7(k + 1), sp.dims, sp.chunk, nullptr);
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
        // operators outside this path (hdf5SCFIntg.F90:563-760 opens them)
        empty_operator_groups(top.id, n, comps, nk);
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
            std::vector<double> cv((size_t)comps * nc * n * nk);
            if (olcao_get_scf_term(ctx, 5, 0, cv.data()) != OLCAO_OK)
                throw H5Error{"reading coreVale from the device"};
            write_core_vale(top.id, cv.data(), n, nc, comps, nk);
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
    if (numStates <= 0 || numStates > n) return OLCAO_ERR_INPUT;
    std::vector<std::vector<double>> w(spin), v(spin);
    for (int q = 0; q < spin; ++q) {
        w[q].resize((size_t)nk * numStates);
        v[q].resize((size_t)nk * numStates * n * comps);
        if ((st = olcao_get_eigen(ctx, q, numStates, w[q].data(), v[q].data())) != OLCAO_OK) return st;
    }
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Handle file{open_or_create(path), H5Fclose};
        write_eigen_groups(file.id, w, v, n, nk, numStates, gamma);
    } catch (const H5Error &e) {
        std::fprintf(stderr, "olcao_write_scf_eigen_h5: %s failed for %s\n", e.what.c_str(), path);
        return OLCAO_ERR_INPUT;
    }
    return OLCAO_OK;
}

extern "C" olcao_status olcao_write_pscf_h5(olcao_ctx *ctx, const double *potCoeffs, int numStates,
                                            const char *kPointName, const char *path)
{
    if (!ctx || !potCoeffs || !kPointName || !*kPointName || !path) return OLCAO_ERR_INPUT;
    long long dims[8];
    int flags[4];
    olcao_status st = olcao_get_dims(ctx, dims);
    if (st == OLCAO_OK) st = olcao_get_flags(ctx, flags);
    if (st != OLCAO_OK) return st;
    const int n = (int)dims[0], nc = (int)dims[1], nk = (int)dims[2];
    const bool gamma = flags[0] != 0;
    const int comps = gamma ? 1 : 2;
    // the reference creates one atomHamOverlap dataset per k and spin under
    // the same name and stops at the second spin
    // (hdf5PSCFIntg.F90:338-344, 'Failed to create hamiltonian overlap')
    if (flags[2] != 1) return OLCAO_ERR_UNSUPPORTED;
    if (numStates <= 0 || numStates > n) return OLCAO_ERR_INPUT;
    // the eigenpairs of the last solve first: the sweeps below rebuild the
    // matrices the solve left
    std::vector<std::vector<double>> w(1), v(1);
    w[0].resize((size_t)nk * numStates);
    v[0].resize((size_t)nk * numStates * n * comps);
    if ((st = olcao_get_eigen(ctx, 0, numStates, w[0].data(), v[0].data())) != OLCAO_OK) return st;
    const size_t full = (size_t)comps * n * n * nk;
    std::vector<double> S(full), H(full);
    std::vector<double> cv(std::max<size_t>((size_t)comps * nc * n * nk, 1)),
        cc(std::max<size_t>((size_t)comps * nc * nc * nk, 1)), cvH(cv.size()), ccH(cc.size());
    if ((st = olcao_gauss_overlap(ctx, 0, S.data(), cv.data(), cc.data())) != OLCAO_OK) return st;
    if ((st = olcao_gauss_overlap_ham_pscf(ctx, potCoeffs, H.data(), cvH.data(), ccH.data())) != OLCAO_OK)
        return st;
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        Handle file{open_or_create(path), H5Fclose};
        if (H5Lexists(file.id, kPointName, H5P_DEFAULT) > 0)
            check(H5Ldelete(file.id, kPointName, H5P_DEFAULT), "old k-point group", 0);
        Handle kg{check(H5Gcreate2(file.id, kPointName, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                        "k-point group"),
                  H5Gclose};
        Handle top{check(H5Gcreate2(kg.id, "atomIntgGroup", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                         "atomIntgGroup"),
                   H5Gclose};
        {
            Handle g{check(H5Gcreate2(top.id, "atomOverlap", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                           "atomOverlap"),
                     H5Gclose};
            write_packed_term(g.id, S.data(), n, comps, nk);
        }
        {
            Handle g{check(H5Gcreate2(top.id, "atomHamOverlap", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                           "atomHamOverlap"),
                     H5Gclose};
            write_packed_term(g.id, H.data(), n, comps, nk);
        }
        if (nc > 0) write_core_vale(top.id, cv.data(), n, nc, comps, nk);
        empty_operator_groups(top.id, n, comps, nk);
        write_eigen_groups(kg.id, w, v, n, nk, numStates, gamma);
    } catch (const H5Error &e) {
        std::fprintf(stderr, "olcao_write_pscf_h5: %s failed for %s\n", e.what.c_str(), path);
        return OLCAO_ERR_INPUT;
    }
    return OLCAO_OK;
}

// gaussKOverlap's output (integrals3Terms.F90:956-1384 with ortho3Terms
// opcodes 8 and 9, :1547-1566): per axis and k-point the full valeDim x
// valeDim matrix as "real<k>" and "imag<k>" datasets of atomKOverlap/<axis>
// and atomKOverlapPlusG/<axis>, each group's status set to 1.  The reference
// passes recipVectors as plusG for the second group but its call of
// KOverlap2CIntg uses recipVectors(:, axis) / numAxialKPoints(axis) alone
// (the plusG term is commented out, :1286-1296), so both groups hold the
// same matrices.
extern "C" olcao_status olcao_write_k_overlap_h5(olcao_ctx *ctx, const int numAxial[3],
                                                 const char *kPointName, const char *path)
{
    if (!ctx || !numAxial || !path) return OLCAO_ERR_INPUT;
    long long dims[8];
    int flags[4];
    olcao_status st = olcao_get_dims(ctx, dims);
    if (st == OLCAO_OK) st = olcao_get_flags(ctx, flags);
    if (st != OLCAO_OK) return st;
    // the Gamma build compiles gaussKOverlap out (OLCAO.F90:303-310)
    if (flags[0]) return OLCAO_ERR_UNSUPPORTED;
    const int n = (int)dims[0], nc = (int)dims[1], nk = (int)dims[2];
    const size_t full = (size_t)2 * n * n * nk;
    std::vector<double> mat(3 * full), part((size_t)n * n);
    std::vector<double> cv(std::max<size_t>(3 * (size_t)2 * nc * n * nk, 1)),
        cc(std::max<size_t>(3 * (size_t)2 * nc * nc * nk, 1));
    if ((st = olcao_gauss_k_overlap(ctx, numAxial, mat.data(), cv.data(), cc.data())) != OLCAO_OK)
        return st;
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        if (H5Fis_hdf5(path) <= 0) throw H5Error{"opening a file that is not HDF5"};
        Handle file{check(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), "file open"), H5Fclose};
        const std::string top = (kPointName && *kPointName ? "/" + std::string(kPointName) : std::string()) +
                                "/atomIntgGroup/";
        for (const char *name : {"atomKOverlap", "atomKOverlapPlusG"}) {
            const std::string grp = top + name;
            Handle g{check(H5Gopen2(file.id, grp.c_str(), H5P_DEFAULT), "k-overlap group"), H5Gclose};
            for (int x = 0; x < 3; ++x)
                for (int k = 0; k < nk; ++k) {
                    const double *m = mat.data() + x * full + (size_t)2 * n * n * k;
                    for (int q = 0; q < 2; ++q) {
                        for (size_t e = 0; e < (size_t)n * n; ++e) part[e] = m[2 * e + q];
                        const std::string ds = grp + "/" + name7(x + 1) + "/" + (q == 0 ? "real" : "imag") +
                                               name7(k + 1);
                        Handle d{check(H5Dopen2(file.id, ds.c_str(), H5P_DEFAULT), "k-overlap dataset"),
                                 H5Dclose};
                        check(H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, part.data()),
                              "k-overlap write", 0);
                    }
                }
            Handle a{check(H5Aopen(g.id, "status", H5P_DEFAULT), "status attribute"), H5Aclose};
            const int one = 1;
            check(H5Awrite(a.id, H5T_NATIVE_INT, &one), "status write", 0);
        }
    } catch (const H5Error &e) {
        std::fprintf(stderr, "olcao_write_k_overlap_h5: %s failed for %s\n", e.what.c_str(), path);
        return OLCAO_ERR_INPUT;
    }
    return OLCAO_OK;
}
