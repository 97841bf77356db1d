/* olcao_gpu.h -- C-ABI of the MI355X (gfx950) integral -> H(k)/S(k) ->
 * eigenvalue path of OLCAO.
 *
 * The reference has no plugin API; its boundary is the set of Fortran module
 * procedures that src/uolcao/OLCAO.F90 calls (setupSCF :248-279, mainSCF
 * :482, intgPSCF :645-662, bandPSCF :758).  Each entry point below replaces
 * one of them; the citation names the procedure.  The Fortran facade
 * (olcao_amd/fortran/olcao_gpu_facade.f90) binds these with ISO_C_BINDING and
 * the Python harness binds them with ctypes.
 *
 * Conventions
 *   - Plain C types only; every pointer is host memory owned by the caller.
 *   - Arrays mirror the reference's Fortran arrays (column-major).  A complex
 *     matrix valeVale(n,n,nk) is nk*n*n interleaved (re,im) doubles, element
 *     (row r, column c, k-point k) at offset 2*(r + c*n + k*n*n).  The Gamma
 *     build (system.gamma = 1) uses real n*n matrices.
 *   - Integral matrices hold the upper triangle (row <= column) as the
 *     reference's valeVale does; the strictly lower part of every diagonal
 *     atom block is zero (intgSaving.F90:484-509).
 *   - Indices are 0-based.
 *   - Every call returns olcao_status; olcao_last_error() gives the message.
 *     The reference stops on the same conditions (integrals.F90:377-384,
 *     interfaces.F90:107-110).
 *   - One context is used by one host thread at a time.  All device state
 *     (system, work lists, S(k), H(k), term matrices) persists in the context
 *     across SCF iterations.
 */
#ifndef OLCAO_GPU_H
#define OLCAO_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct olcao_ctx olcao_ctx;

typedef enum {
    OLCAO_OK = 0,
    OLCAO_ERR_HIP = 1,          /* HIP runtime failure */
    OLCAO_ERR_ALLOC = 2,        /* device allocation failed */
    OLCAO_ERR_INPUT = 3,        /* invalid system description */
    OLCAO_ERR_CELLS = 4,        /* cell list too small (integrals.F90:377) */
    OLCAO_ERR_NOT_SPD = 5,      /* S(k) not positive definite (zhegv info>n) */
    OLCAO_ERR_SOLVER = 6,       /* eigensolver failed to converge */
    OLCAO_ERR_UNSUPPORTED = 7,  /* option not available in this build */
    OLCAO_ERR_STATE = 8,        /* call order violated (e.g. solve before build) */
    OLCAO_ERR_NO_DEVICE = 9     /* no GPU visible */
} olcao_status;

/* Arithmetic of the integral build.  OLCAO_PREC_FP64: everything in FP64
 * (the drop-in form).  OLCAO_PREC_MIXED: the Bloch sums as exact products
 * of BF16 slices on the matrix cores with FP32 accumulation (a relative
 * error of about 2^-48 of each block column's largest entry per cell), and the
 * three-Gaussian exponentials of distant potential images (those below
 * 1e-4) by the FP32 hardware exponential after an FP64 range reduction;
 * everything else as FP64.  The screening exponents, the Cholesky
 * factorisation and the eigensolve are FP64 in both forms. */
enum { OLCAO_PREC_FP64 = 0, OLCAO_PREC_MIXED = 1 };

/* The system as the reference holds it after parseInput
 * (input.f90:116-244) and before any derived setup. */
typedef struct {
    int nAtoms;
    const double *atomPos;        /* [nAtoms][3] Cartesian bohr */
    const int *atomType;          /* [nAtoms] 0-based */
    int nTypes, maxAlphas, maxRadial;
    const int *typeNumAlphas;     /* [nTypes] */
    const double *typeAlphas;     /* [nTypes][maxAlphas] strictly ascending */
    const int *typeNumOrbAlphas;  /* [nTypes][4] alphas used by l=0..3 */
    const int *typeElement;       /* [nTypes] */
    const int *typeNumValeRadial; /* [nTypes] */
    const int *typeNumCoreRadial; /* [nTypes] */
    const int *typeValeL;         /* [nTypes][maxRadial] */
    const int *typeCoreL;         /* [nTypes][maxRadial] */
    const double *typeValeCoef;   /* [nTypes][maxRadial][maxAlphas] */
    const double *typeCoreCoef;   /* [nTypes][maxRadial][maxAlphas] */
    int nPotTypes, maxPotAlphas;
    const int *potNumAlphas;      /* [nPotTypes] */
    const double *potAlphas;      /* [nPotTypes][maxPotAlphas] ascending */
    const double *potZ;           /* [nPotTypes] nuclear charge */
    const double *potNucAlpha;    /* [nPotTypes] */
    const int *potElement;        /* [nPotTypes] */
    int nPotSites;
    const double *potPos;         /* [nPotSites][3], grouped by type */
    const int *potSiteType;       /* [nPotSites] 0-based */
    double cellVectors[9];        /* a_i = cellVectors[3*i + xyz], bohr */
    int nk;
    const double *kpoints;        /* [nk][3] fractional reciprocal coords */
    double basisCutoff;           /* BASISFUNCTION cutoff, e.g. 1e-16 */
    int gamma;                    /* 1: real Gamma-point build */
    int rel;                      /* 1: scalar relativistic, adds the mass-velocity
                                     term (O_Potential rel, potential.f90) */
    int spin;                     /* 1 or 2 potential-coefficient sets */
} olcao_system;

/* Generalised eigensolve: OLCAO_EIG_AUTO uses the batched in-LDS solver
 * (one workgroup per k-point: Cholesky, reduction, Jacobi) where the matrices
 * fit (complex n <= 70, real n <= 99), else the device tridiagonal path
 * (blocked Cholesky, Householder reduction, divide and conquer, blocked
 * back-transform): its reduction keeps the matrix in LDS across a group of
 * workgroups up to n = 896 and works on it in global memory by panels of
 * 32 columns beyond (real n <= 7168, complex n <= 4096); rocSOLVER past
 * those sizes.  OLCAO_EIG_LIBRARY always uses rocSOLVER; OLCAO_EIG_BLOCKED
 * takes the global-memory reduction also where the LDS one fits.  The
 * drop-in solvers (olcao_solve_zhegv/dsygv) read the same choice from the
 * environment variable OLCAO_EIGENSOLVER (auto, library or blocked). */
enum { OLCAO_EIG_AUTO = 0, OLCAO_EIG_LIBRARY = 1, OLCAO_EIG_BLOCKED = 2 };

/* The tile shapes chunkItems, segPairs and bloch below are what the
 * project's autotuner (olcao_amd/autotune.py) chooses per workload.
 * Bloch-sum kernel: OLCAO_BLOCH_AUTO and OLCAO_BLOCH_MFMA take the
 * matrix-core kernel (FP64 MFMA in the FP64 form, exact split BF16 products
 * in the mixed form); OLCAO_BLOCH_MFMA64 the FP64 MFMA kernel in both
 * forms; OLCAO_BLOCH_VALU the vector-ALU kernel (FP64 in both forms). */
enum { OLCAO_BLOCH_AUTO = 0, OLCAO_BLOCH_VALU = 1, OLCAO_BLOCH_MFMA = 2, OLCAO_BLOCH_MFMA64 = 3 };

typedef struct {
    int precision;                /* OLCAO_PREC_* */
    int eigensolver;              /* OLCAO_EIG_* */
    int bloch;                    /* OLCAO_BLOCH_* */
    int graph;                    /* 1: olcao_scf_iteration replays a captured graph
                                     of its device work (build and solve) after its
                                     first call; the library eigensolver (real n >
                                     7168, complex n > 4096) is not captured and
                                     such contexts stay eager */
    int chunkItems;               /* three-centre work per lane: image-list chunks of
                                     about this many items (16..1048576; 0: 192) */
    int segPairs;                 /* contraction segment length in primitive pairs
                                     (1..64; 0: 32) */
    int reserved[2];
} olcao_options;

/* Context lifetime.  Replaces the derived setup of OLCAO.F90:setupSCF
 * (:196-244): initializeLattice (lattice.f90:310), initializeFindVec (:892),
 * initializeKPoints (kpoints.f90:937), renormalizeBasis (basis.f90:55),
 * makeAlphaDist (gaussRelations.f90:64), plus the screened work lists.
 *
 * Threads: one context is used by one host thread at a time.  Several
 * contexts are not to be iterated from concurrent host threads outside
 * olcao_scf_iteration_batch: with rocSOLVER solves (n > 640) overlapping
 * other contexts' kernels, wrong eigenvalues were observed once for one or
 * two of 19 such cells, with the cause not found; the batched call runs
 * those solves alone. */
olcao_status olcao_ctx_create(const olcao_system *sys, const olcao_options *opt,
                              olcao_ctx **out);
void olcao_ctx_destroy(olcao_ctx *ctx);
const char *olcao_last_error(const olcao_ctx *ctx);

/* dims[0..7] = valeDim, coreDim, nk, potDim, nCells, nPairCells,
 * nPrimitivePairs, nAtomPairs. */
olcao_status olcao_get_dims(const olcao_ctx *ctx, long long dims[8]);

/* Eigenpairs of the last solve for one spin (0-based), as the eigenvalue
 * and eigenvector datasets of secularEqnSCF hold them (secularEqn.F90:
 * 259-313): evals [nk][numStates] ascending, evecs [nk][numStates][valeDim]
 * complex (real at Gamma).  Either pointer may be NULL. */
olcao_status olcao_get_eigen(const olcao_ctx *ctx, int spin, int numStates, double *evals,
                             double *evecs);

/* flags[0..3] = gamma, rel, spin, precision of the context. */
olcao_status olcao_get_flags(const olcao_ctx *ctx, int flags[4]);

/* Derived setup for inspection: the sorted cell list (cellDimsReal,
 * cellSizesReal, lattice.f90:571-695) and the Cartesian k-points
 * (kpoints.f90:424-454).  Any pointer may be NULL. */
olcao_status olcao_get_setup(const olcao_ctx *ctx, double *cellDims,
                             double *cellSizes, double *kCart);

/* Two-centre integral sweeps.  op: 0 overlap (gaussOverlapOL,
 * integrals.F90:151), 1 kinetic (gaussOverlapKE, :656), 2 mass velocity
 * (gaussOverlapMV, :1091; needs rel = 1; stored negated, as the term enters
 * H).  The result stays on the device (S(k) for op 0) and is copied to
 * vv/cv/cc when non-NULL.  vv is valeVale as the reference writes it: when
 * coreDim > 0 it is orthogonalised against the core (orthoOL / ortho,
 * integrals.F90:3608-3910) with components below 1e-10 zeroed
 * (intgOrtho.F90:142-190); cv and cc are the raw coreVale(coreDim,valeDim,nk)
 * and coreCore(coreDim,coreDim,nk).  Operator sweeps run the overlap sweep
 * first when coreDim > 0 and it has not run. */
olcao_status olcao_gauss_overlap(olcao_ctx *ctx, int op, double *vv,
                                 double *cv, double *cc);

/* Nuclear-attraction sweep (gaussOverlapNP, integrals.F90:1531 with
 * nuclearPE :2533). */
olcao_status olcao_gauss_overlap_np(olcao_ctx *ctx, double *vv, double *cv,
                                    double *cc);

/* One electronic-potential term matrix V_j (gaussOverlapEP,
 * integrals.F90:1946 with electronicPE :2700, called per term by
 * elecPotGaussOverlap :2357).  term: 0-based over potDim, type-major. */
olcao_status olcao_elec_pot_gauss_overlap(olcao_ctx *ctx, int term, double *vv,
                                          double *cv, double *cc);

/* Dipole-moment matrices (gaussOverlapDM, integrals3Terms.F90:64, with
 * dipole3CIntg gaussIntegrals.f90:49404): the three Cartesian components of
 * <phi_i| r - center |phi_j(. - L)> Bloch-summed like the overlap, each
 * orthogonalised against the core (ortho3Terms opCode 6, :1388) and stored
 * upper-triangular as valeVale.  center: the dipole centre (O_Input
 * dipoleCenter), bohr.  vv, cv, cc: three blocks of the layout of
 * olcao_gauss_overlap's, component-major (x, y, z). */
olcao_status olcao_gauss_overlap_dm(olcao_ctx *ctx, const double *center, double *vv,
                                    double *cv, double *cc);

/* Momentum matrices (gaussOverlapMM, integrals3Terms.F90:511, with
 * momentum2CIntg gaussIntegrals.f90:23041): the gradient integrals
 * <phi_i| d/dr |phi_j(. - L)> Bloch-summed with the factor -i
 * (applyPhaseFactors runCode 1, intgSaving.F90:178-187; -X at Gamma,
 * :727-731), i.e. the Hermitian matrices of p = -i nabla; orthogonalised
 * (opCode 7) and laid out as olcao_gauss_overlap_dm. */
olcao_status olcao_gauss_overlap_mm(olcao_ctx *ctx, double *vv, double *cv, double *cc);

/* gaussKOverlap (integrals3Terms.F90:956-1384): the overlap with
 * exp(-i deltaK . r), deltaK = recipVectors(:, axis) / numAxialKPoints(axis),
 * for the three axes, orthogonalised against the core as ortho3Terms opcode
 * 8 does (full matrix, no zeroing of small entries; integrals3Terms.F90
 * :1388-1608).  numAxial: the k-mesh divisions per axis.  vv [3][nk][n][n],
 * cv [3][nk][nc][n], cc [3][nk][nc][nc] complex, column-major per k (full
 * matrices: the pair blocks i <= j as saveCurrentPair places them, then the
 * orthogonalisation over the whole matrix).  The reference's
 * KOverlapPlusG group holds the same integrals (its plusG argument selects
 * only the output group).  Complex contexts only: a Gamma context returns
 * OLCAO_ERR_UNSUPPORTED. */
olcao_status olcao_gauss_k_overlap(olcao_ctx *ctx, const int numAxial[3], double *vv, double *cv,
                                   double *cc);

/* computeForceIntg (forces.F90:60-898): the force-derivative matrices for
 * the potential coefficients coef [spin][potDim] and totalEnergy.  Per spin
 * q and Cartesian component x (slot 3 q + x), the Bloch sums over the fused
 * Hamiltonian sweep's screened images of the derivatives with respect to
 * the ket centre B_x (the potential and nuclear centres held fixed) of
 *   - sum_j coef_j (three-Gaussian block j)   (delectron3CIntgCB, :581)
 *   + Z (Gaussian-damped nuclear block)        (dnuclear3CIntgCB, :611)
 *   - kinetic block                            (dkinetic2CIntg, :648)
 * plus -2 totalEnergy times the momentum block (momentum2CIntg, :715).
 * vv [3 spin][nk][n][n], cv [3 spin][nk][nc][n], cc [3 spin][nk][nc][nc],
 * complex column-major per k (real at Gamma), not orthogonalised.  A
 * complex build saves each atom pair in full (saveCurrentPair saveFull = 1:
 * the block and its dagger, a diagonal atom block replaced by its own
 * dagger); at Gamma the upper triangle as the other sweeps
 * (saveCurrentPairGamma).  Bases with f orbitals return
 * OLCAO_ERR_UNSUPPORTED: the f-f nuclear derivative needs F_7, which the
 * reference's Boys function does not give (its dnuclear3CIntgCB is off by
 * about 1e-2 there). */
olcao_status olcao_force_integrals(olcao_ctx *ctx, const double *coef, double totalEnergy,
                                   double *vv, double *cv, double *cc);

/* Fused Hamiltonian sweep with the potential coefficients folded in
 * (gaussOverlapHamPSCF, integrals.F90:2860): KE (- MV when rel = 1)
 * + sum_j c_j V_j - sum Z N, orthogonalised when coreDim > 0
 * (orthoHamPSCF :3913).  potCoeffs: [spin][potDim]; vv, cv, cc hold one
 * block per spin, spin-major. */
olcao_status olcao_gauss_overlap_ham_pscf(olcao_ctx *ctx, const double *potCoeffs,
                                          double *vv, double *cv, double *cc);

/* Generalised eigensolve of the device-resident H(k), S(k) from the two
 * calls above, for all k and spins (secularEqnPSCF, secularEqn.F90:331, with
 * solveZHEGV interfaces.F90:73 / solveDSYGV :152).  evals:
 * [spin][nk][numStates] ascending; evecs (may be NULL):
 * [spin][nk][numStates][valeDim] complex (column s of eigenvector k at
 * 2*(r + s*n + k*n*numStates) within a spin), real at Gamma. */
olcao_status olcao_secular_eqn_pscf(olcao_ctx *ctx, int numStates, double *evals,
                                    double *evecs);

/* Mode A setup: build and keep S, KE, NP, MV (rel = 1) and every V_j on
 * the device, each orthogonalised when coreDim > 0 (setupSCF,
 * OLCAO.F90:248-279). */
olcao_status olcao_setup_scf(olcao_ctx *ctx);

/* The device-resident matrices olcao_setup_scf built, copied to host, in the
 * layout of olcao_gauss_overlap's vv (cv for which = 5): which 0 overlap
 * (atomOverlap), 1 kinetic (atomKEOverlap), 2 mass velocity (atomMVOverlap,
 * rel = 1), 3 nuclear (atomNPOverlap), 4 potential term `term`
 * (atomPotOverlap), 5 raw coreVale overlap (atomOverlapCV) -- the
 * matrices setupSCF leaves in its integral file (hdf5SCFIntg.F90:188-460). */
olcao_status olcao_get_scf_term(const olcao_ctx *ctx, int which, int term, double *out);

/* Mode A iteration: H = NP + KE (+ MV) + sum_j c_j V_j per spin (MV and a
 * coefficient of exactly 0 add their term unscaled, matrixSubs.F90:369-373),
 * then the solve for all k (secularEqnSCF, secularEqn.F90:37).  potCoeffs:
 * [spin][potDim]; evals, evecs as olcao_secular_eqn_pscf. */
olcao_status olcao_secular_eqn_scf(olcao_ctx *ctx, const double *potCoeffs,
                                   int numStates, double *evals, double *evecs);

/* Valence density matrix and its traces (makeValenceRho with inSCF = 1,
 * valeCharge.F90:34-526): rho(k) = sum_j f_jk x_jk x_jk^H per spin from the
 * eigenvectors of the last solve (:214-232; dsyr at Gamma :266-277), then
 * Tr[M rho] summed over k as matrixElementMult does on packed matrices
 * (matrixSubs.F90:66-180) for M = S, NP, KE, MV and every V_j (:330-418).
 * At spin = 2 the density is first rewritten as up + down and up - down
 * (:306-314).  Needs olcao_setup_scf (the term matrices) and a solve.
 * occupancy: [nk][spin][numStates], electronPopulation's order (:160-169);
 * a state whose populations sum below 1e-8 in magnitude is skipped.
 * electronEnergy: [spin] sum f * eigenvalue per spin; traces: [spin][4] =
 * chargeDensityTrace, nucPotTrace, kineticEnergyTrace, massVelocityTrace
 * (0 when rel = 0); potRho: [spin][potDim]; index 0 is the total density and
 * index 1 (spin = 2) the difference.  Any output pointer may be NULL. */
olcao_status olcao_vale_charge(olcao_ctx *ctx, int numStates, const double *occupancy,
                               double *electronEnergy, double *traces, double *potRho);

/* Mode B iteration, the unit the benchmark times: overlap sweep, fused
 * Hamiltonian sweep per spin and the solve for all k (intgPSCF + bandPSCF,
 * OLCAO.F90:533-770).  potCoeffs: [spin][potDim]; evals:
 * [spin][nk][numStates]. */
olcao_status olcao_scf_iteration(olcao_ctx *ctx, const double *potCoeffs,
                                 int numStates, double *evals);

/* Batched Mode-B iteration over independent cells (contexts), the batched
 * form for many uneven cells.  It has no counterpart in the reference, which
 * runs one cell per process through OLCAO.F90:533-770; each cell's result is
 * what olcao_scf_iteration gives for it.  nWorkers host threads (1..16) take
 * the cells heaviest first, each cell on its own streams, so the device runs
 * several cells at once: every cell's integral build, followed by its solve
 * where that is the one-workgroup LDS solver; the other solves (the
 * tridiagonal path, a cooperative launch, and rocSOLVER's) run last with
 * nothing else on the device, the k-points of all cells with matrices of
 * the same size and kind in one batched solve.
 * potCoeffs[i], numStates[i], evals[i] as olcao_scf_iteration for cell i.
 * cellMs (may be NULL): [nCells] each cell's time from its first to its last
 * event within the batch (its latency while sharing the device).  On failure
 * *failedCell names the cell and olcao_last_error(NULL) its message.  The
 * contexts must be distinct and not used by another thread meanwhile. */
olcao_status olcao_scf_iteration_batch(int nCells, olcao_ctx *const *ctxs,
                                       const double *const *potCoeffs, const int *numStates,
                                       double *const *evals, int nWorkers, double *cellMs,
                                       int *failedCell);

/* The batched iteration with a choice of how the cells' integral builds
 * share the device (no counterpart in the reference).
 * OLCAO_BATCH_CELLS: as olcao_scf_iteration_batch, every cell's whole build
 * on its own streams, concurrently from the worker threads.
 * OLCAO_BATCH_SHARED: per spin, the cells' own stages (overlap sweep,
 * kinetic blocks; afterwards contraction, Bloch sums, orthogonalisation)
 * from the worker threads, and between them the fused Hamiltonian sweep's
 * three-centre work of all cells in shared launches, one per angular class
 * and group of cells, waves taking 64 chunks of one cell from a task list
 * that interleaves the cells' chunk lists heaviest first.
 * Each cell's result is the same in both modes; other arguments as
 * olcao_scf_iteration_batch. */
enum { OLCAO_BATCH_CELLS = 0, OLCAO_BATCH_SHARED = 1 };
olcao_status olcao_scf_iteration_batch_mode(int nCells, olcao_ctx *const *ctxs,
                                            const double *const *potCoeffs, const int *numStates,
                                            double *const *evals, int nWorkers, int mode,
                                            double *cellMs, int *failedCell);

/* Stage times (ms, HIP events) of the last olcao_scf_iteration:
 * t[0] overlap sweep (incl. its contraction and Bloch sum), t[1] Hamiltonian
 * primitive blocks, t[2] Hamiltonian contraction + Bloch sum (+ ortho),
 * t[3] solve, t[4] total; t[5] three-centre kernels alone (all summed over
 * spins); t[6] 1 when the iteration was a graph replay (then only t[4] is
 * measured, from the coefficient upload to the graph's end); t[7] the
 * tridiagonal reduction kernel (k_trd) alone, summed over spins (0 when the
 * solve took another path). */
olcao_status olcao_get_timings(const olcao_ctx *ctx, double t[8]);

/* Graph mode of a context (olcao_options.graph): *state 0 off, 1 a captured
 * iteration graph is current, 2 the next olcao_scf_iteration runs eagerly
 * and captures (first call, or a buffer was reallocated since the capture),
 * 3 capture refused; *reason (may be NULL) then points to the context's
 * message, valid until the context is destroyed. */
olcao_status olcao_get_graph_state(const olcao_ctx *ctx, int *state, const char **reason);

/* Screened primitive pairs per angular class k = 4 la + lb (the finish
 * kernels' work units). */
olcao_status olcao_get_class_pairs(const olcao_ctx *ctx, long long counts[16]);

/* Work counters of the last three-centre sweep: counts[2*(4*la+lb)] three-
 * Gaussian overlap evaluations and counts[2*(4*la+lb)+1] nuclear
 * evaluations of angular class (la, lb). */
olcao_status olcao_get_work_counts(olcao_ctx *ctx, unsigned long long counts[32]);

/* Drop-in generalised Hermitian solver for host matrices, batched:
 * replaces solveZHEGV (interfaces.F90:73-120): A (n,n) upper triangle in,
 * eigenvectors out; B (n,n) upper triangle in.  w: [nbatch][n]. */
olcao_status olcao_solve_zhegv(int n, int nbatch, double *A, double *B, double *w);
/* Real twin, replaces solveDSYGV (interfaces.F90:152-200). */
olcao_status olcao_solve_dsygv(int n, int nbatch, double *A, double *B, double *w);

/* olcao_solve_zhegv without device allocation.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as olcao_solve_zhegv (host buffers), for n <= 70, the batched
 * in-LDS solver's range; the batch is staged in a fixed 64 MiB device array
 * per GPU (at most 64 MiB / ((4 n^2 + n + 1) * 8 B) matrices per call,
 * e.g. 645 at n = 57), larger batches are refused with OLCAO_ERR_INPUT.
 * The array is shared by every call on a GPU and keeps the last call's
 * contents: calls that overlap on one GPU (from several host threads)
 * race.  Runs on the null stream and synchronises with the host before it
 * returns; the message of a failure is olcao_last_error(NULL). */
olcao_status olcao_solve_zhegv_noalloc(int n, int nbatch, double *A, double *B, double *w);

/* Primitive integral blocks on the device, for parity tests of the kernels'
 * device functions against the reference subroutines overlap2CIntg
 * (gaussIntegrals.f90:26, op 0), kinetic2CIntg (:4422, op 1),
 * electron3CIntg (:11624, op 2), nuclear3CIntg (:16223, op 3),
 * massvel2CIntg (:39070, op 4), and the real (op 7) and imaginary (op 8)
 * parts of KOverlap2CIntg (:129408) with deltaK = C.
 * a: [n][3] exponents, A/B/C: [n][3], sw: [n] l1l2switch, sh: [n][16][16]
 * column-major blocks (sh(i,j) at i + 16 j); entries outside the class block
 * are zero. */
olcao_status olcao_prim_batch(int op, int n, const double *a, const double *A,
                              const double *B, const double *C, const int *sw,
                              double *sh);

/* Boys function F_0..F_6 on the device (gaussIntegrals.f90:16080).
 * T: [n], F: [n][7]. */
olcao_status olcao_boys_batch(int n, const double *T, double *F);

#ifdef __cplusplus
}
#endif
#endif
