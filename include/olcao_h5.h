/* olcao_h5.h -- the setup's integral file in the reference's HDF5 layout.
 *
 * libolcao_h5.so (host code over the HDF5 C library) writes the matrices
 * that olcao_setup_scf keeps on the device into the groups and datasets that
 * setupSCF writes (src/uolcao/hdf5SCFIntg.F90:105-460, written by
 * integrals.F90's gaussOverlap* / elecPotGaussOverlap and ortho*), so a
 * reference reader (secularEqnSCF's readPackedMatrix, valeCharge) finds them
 * where it looks:
 *
 *   /atomIntgGroup/atomOverlap/<k>          packed S(k)      (1 or 2, n(n+1)/2)
 *   /atomIntgGroup/atomKEOverlap/<k>        packed KE(k)
 *   /atomIntgGroup/atomMVOverlap/<k>        packed MV(k)     (rel = 1)
 *   /atomIntgGroup/atomNPOverlap/<k>        packed NP(k)
 *   /atomIntgGroup/atomPotOverlap/<j>/<k>   packed V_j(k)
 *   /atomIntgGroup/atomOverlapCV/real<k>, imag<k>   coreVale (coreDim, valeDim)
 *
 * <k> and <j> are 1-based, written "%07d".  Dataset dimensions are given in
 * the reference's Fortran order; components are (re, im), one at Gamma.
 * Packed index i(i-1)/2 + j for column i, row j <= i (matrixSubs.F90:506).
 * Datasets are chunked (at most 250M elements per chunk, hdf5SCFIntg.F90:
 * 144-175) and deflated at level 1.  Every group that holds a term gets the
 * integer attribute "status" = 1 once written, as the reference records a
 * finished calculation (integrals.F90:3728-3731).  The operator groups
 * (atomDMOverlap, atomMMOverlap, atomKOverlap, atomKOverlapPlusG) are
 * created with status 0 and filled by olcao_write_scf_vector_terms_h5 and
 * olcao_write_k_overlap_h5. */
#ifndef OLCAO_H5_H
#define OLCAO_H5_H

#include "olcao_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Write the file at `path` (truncating it) from a context after
 * olcao_setup_scf.  Returns OLCAO_ERR_STATE before the setup and
 * OLCAO_ERR_INPUT when the file cannot be written. */
olcao_status olcao_write_scf_integrals_h5(olcao_ctx *ctx, const char *path);

/* Fill the dipole-moment (which bit 0, about `dipoleCenter`) and momentum
 * (which bit 1) groups of an integral file written by
 * olcao_write_scf_integrals_h5: computes them on the device
 * (olcao_gauss_overlap_dm / _mm), writes each component's packed matrices
 * into /atomIntgGroup/atomDMOverlap/<x>/<k> (atomMMOverlap likewise, x =
 * 0000001..0000003) and sets the group's status to 1, as gaussOverlapDM /
 * gaussOverlapMM with ortho3Terms leave the file when setupSCF computes them
 * for properties (OLCAO.F90:290-302, integrals3Terms.F90:1388-1608).
 * dipoleCenter may be NULL when bit 0 is clear.  OLCAO_ERR_INPUT when the
 * file is not such an integral file. */
olcao_status olcao_write_scf_vector_terms_h5(olcao_ctx *ctx, int which,
                                             const double *dipoleCenter, const char *path);

/* Add the eigenpairs of the last solve to the HDF5 file at `path` (created
 * when absent, OLCAO_ERR_INPUT when the path holds another kind of file;
 * groups of an earlier call are replaced) as secularEqnSCF writes
 * them (hdf5SCFEigVal.f90:60-95, hdf5SCFEigVec.F90:70-160,
 * secularEqn.F90:259-313):
 *
 *   /eigenValues/<k><s>                      (numStates), ascending, Hartree
 *   /eigenVectors/real<k><s>, imag<k><s>     (valeDim, numStates)
 *   /eigenVectors/<k><s>                     the same at Gamma (real)
 *
 * <k><s> = k-point and spin, each "%07d" and 1-based.  Chunked (at most
 * 250M elements per chunk), deflate level 1; each first eigenvector dataset
 * carries the integer attribute "status", 0 as the reference leaves it once
 * all k-points of a spin are done (secularEqn.F90:307-313). */
olcao_status olcao_write_scf_eigen_h5(olcao_ctx *ctx, int numStates, const char *path);

/* The PSCF file of one k-point set (hdf5PSCF.F90:60-190 with
 * hdf5PSCFIntg.F90:95-452, hdf5PSCFEigVal.f90, hdf5PSCFEigVec.F90), as
 * intgPSCF and bandPSCF leave it (OLCAO.F90:533-770), under the group
 * kPointName (the reference names it from the k-point mesh, e.g.
 * "00004_00004_00004"; an existing group of that name is replaced, the
 * file is created when absent, OLCAO_ERR_INPUT when the path holds another
 * kind of file):
 *
 *   <kPointName>/atomIntgGroup/atomOverlap/<k>      S(k), packed, status 1
 *   <kPointName>/atomIntgGroup/atomHamOverlap/<k>   H(k) of potCoeffs, status 1
 *   <kPointName>/atomIntgGroup/atomOverlapCV/{real,imag}<k>   (coreDim > 0)
 *   <kPointName>/atomIntgGroup/atom{DM,MM,K,KPlusG}Overlap/...  empty, status 0
 *   <kPointName>/eigenValues/<k><s>, eigenVectors/...  the last solve's
 *
 * S and H are rebuilt by the sweeps (orthogonalised when coreDim > 0, as
 * gaussOverlapOL and gaussOverlapHamPSCF write them); the eigenpairs are
 * those of the last solve (olcao_secular_eqn_pscf or olcao_scf_iteration),
 * OLCAO_ERR_STATE without one.  The file-level mesh groups of
 * hdf5PSCFField.F90 (wavGroup, rhoGroup, potGroup, mesh) are not written.
 * Spin 2 is refused (OLCAO_ERR_UNSUPPORTED): the reference creates both
 * spins' Hamiltonian datasets under one name and stops at the second
 * (hdf5PSCFIntg.F90:338-344). */
olcao_status olcao_write_pscf_h5(olcao_ctx *ctx, const double *potCoeffs, int numStates,
                                 const char *kPointName, const char *path);

/* Fill the k-overlap groups of a file written by olcao_write_scf_integrals_h5
 * (kPointName NULL or empty) or of the k-point group kPointName of a file
 * written by olcao_write_pscf_h5, as gaussKOverlap with ortho3Terms leaves
 * them when setupSCF or intgPSCF run with doMTOP = 1 (OLCAO.F90:302-310,
 * :689-698; integrals3Terms.F90:956-1384, :1547-1566):
 *
 *   .../atomIntgGroup/atomKOverlap/<x>/real<k>, imag<k>       (valeDim, valeDim)
 *   .../atomIntgGroup/atomKOverlapPlusG/<x>/real<k>, imag<k>  the same values
 *
 * x = 0000001..0000003 is the axis of deltaK = recipVectors(:, x) /
 * numAxial[x]; each matrix is the full orthogonalised one of
 * olcao_gauss_k_overlap; both groups' status is set to 1.  The reference's
 * call of KOverlap2CIntg leaves out its plusG argument (the term is commented
 * out, integrals3Terms.F90:1286-1296), so the PlusG group repeats the
 * k-overlap.  OLCAO_ERR_UNSUPPORTED for a Gamma context (the Gamma build has
 * no k-overlap), OLCAO_ERR_INPUT when the file or group is not there. */
olcao_status olcao_write_k_overlap_h5(olcao_ctx *ctx, const int numAxial[3], const char *kPointName,
                                      const char *path);

#ifdef __cplusplus
}
#endif
#endif
