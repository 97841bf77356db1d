! Fortran host facade of the MI355X integral -> eigenvalue path.
!
! Keeps the reference's procedure NAMES, so that the call sites in
! src/uolcao/OLCAO.F90 (setupSCF :248-279, mainSCF :482, intgPSCF :645-662,
! bandPSCF :758) and the solver call sites in src/uolcao/secularEqn.F90
! (:227, :509) call a procedure of the same name; each forwards to the C-ABI
! of include/olcao_gpu.h through ISO_C_BINDING.  The argument lists differ
! from the reference's where the reference passes HDF5 dataset ids and reads
! module globals; the facade takes and returns arrays instead:
!   gaussOverlapOL/KE/MV/NP(valeVale, coreVale, coreCore)
!       reference (numComponents, fullCVDims, packedVVDims, did, CVdid, aid)
!       integrals.F90:151, :656, :1091, :1531: the matrices come back as
!       arrays (valeVale orthogonalised when coreDim > 0); the caller packs
!       and writes them (or calls olcao_write_scf_integrals_h5)
!   elecPotGaussOverlap(term, valeVale, coreVale, coreCore)
!       reference (packedVVDims, did(nk,potDim), aid(potDim)) :2357: one
!       potential term per call instead of all of them
!   gaussOverlapHamPSCF(potCoeffs, valeVale, coreVale, coreCore)
!       reference (did(nk,spin), aid) :2860: the coefficients come in as an
!       argument instead of O_Potential's potCoeffs
!   secularEqnSCF(potCoeffs, numStates, energyEigenValues)
!       reference (spinDirection, numStates) secularEqn.F90:37: all spins in
!       one call, the terms from the device instead of the HDF5 file
!   secularEqnPSCF(numStates, energyEigenValues)
!       reference (spin, numStates, numComponents, ol_did, ham_did, eVal_did,
!       eVec_did, eVec_aid) :331: the matrices stay on the device
!   setupSCF(), makeValenceRho(...): no single reference procedure; they
!       stand for the setup's term set and valeCharge.F90:34
!   solveZHEGV, solveDSYGV: the reference's dummy arguments unchanged
!       (interfaces.F90:73, :152)
! tests/fortran/facade_driver.f90 is a Fortran program that calls these at
! the reference driver's call sites.
!
! Errors follow the reference: the library's message is written to unit 20
! and execution stops (integrals.F90:377-384, interfaces.F90:107-110).
!
! Matrices: complex(double) valeVale(valeDim,valeDim,numKPoints) etc. have
! the interleaved column-major layout the C-ABI uses, so they pass by
! address.  Index arrays are 1-based here and converted to the C-ABI's
! 0-based convention in gpuInit.
module O_GPUFacade
   use iso_c_binding
   implicit none
   private

   integer, parameter :: double = kind(0.0d0)

   public :: gpuInit, gpuFinalize, gpuDims
   public :: gaussOverlapOL, gaussOverlapKE, gaussOverlapMV, gaussOverlapNP, elecPotGaussOverlap
   public :: gaussOverlapHamPSCF, secularEqnPSCF, setupSCF, secularEqnSCF
   public :: solveZHEGV, solveDSYGV

   type, bind(C) :: olcao_system_t
      integer(c_int) :: nAtoms
      type(c_ptr) :: atomPos, atomType
      integer(c_int) :: nTypes, maxAlphas, maxRadial
      type(c_ptr) :: typeNumAlphas, typeAlphas, typeNumOrbAlphas, typeElement
      type(c_ptr) :: typeNumValeRadial, typeNumCoreRadial, typeValeL, typeCoreL
      type(c_ptr) :: typeValeCoef, typeCoreCoef
      integer(c_int) :: nPotTypes, maxPotAlphas
      type(c_ptr) :: potNumAlphas, potAlphas, potZ, potNucAlpha, potElement
      integer(c_int) :: nPotSites
      type(c_ptr) :: potPos, potSiteType
      real(c_double) :: cellVectors(9)
      integer(c_int) :: nk
      type(c_ptr) :: kpoints
      real(c_double) :: basisCutoff
      integer(c_int) :: gamma
      integer(c_int) :: rel
      integer(c_int) :: spin
   end type olcao_system_t

   type, bind(C) :: olcao_options_t
      integer(c_int) :: precision
      integer(c_int) :: eigensolver
      integer(c_int) :: bloch
      integer(c_int) :: graph
      integer(c_int) :: chunkItems
      integer(c_int) :: segPairs
      integer(c_int) :: reserved(2)
   end type olcao_options_t

   type(c_ptr), save :: ctx = c_null_ptr

   interface
      integer(c_int) function c_ctx_create(sys, opt, out) bind(C, name="olcao_ctx_create")
         import :: c_int, c_ptr, olcao_system_t, olcao_options_t
         type(olcao_system_t), intent(in) :: sys
         type(olcao_options_t), intent(in) :: opt
         type(c_ptr), intent(out) :: out
      end function
      subroutine c_ctx_destroy(c) bind(C, name="olcao_ctx_destroy")
         import :: c_ptr
         type(c_ptr), value :: c
      end subroutine
      type(c_ptr) function c_last_error(c) bind(C, name="olcao_last_error")
         import :: c_ptr
         type(c_ptr), value :: c
      end function
      integer(c_int) function c_get_dims(c, dims) bind(C, name="olcao_get_dims")
         import :: c_int, c_ptr, c_long_long
         type(c_ptr), value :: c
         integer(c_long_long), intent(out) :: dims(8)
      end function
      integer(c_int) function c_overlap(c, op, vv, cv, cc) bind(C, name="olcao_gauss_overlap")
         import :: c_int, c_ptr
         type(c_ptr), value :: c, vv, cv, cc
         integer(c_int), value :: op
      end function
      integer(c_int) function c_np(c, vv, cv, cc) bind(C, name="olcao_gauss_overlap_np")
         import :: c_int, c_ptr
         type(c_ptr), value :: c, vv, cv, cc
      end function
      integer(c_int) function c_ep(c, term, vv, cv, cc) bind(C, name="olcao_elec_pot_gauss_overlap")
         import :: c_int, c_ptr
         type(c_ptr), value :: c, vv, cv, cc
         integer(c_int), value :: term
      end function
      integer(c_int) function c_ham(c, coef, vv, cv, cc) bind(C, name="olcao_gauss_overlap_ham_pscf")
         import :: c_int, c_ptr
         type(c_ptr), value :: c, coef, vv, cv, cc
      end function
      integer(c_int) function c_sec_pscf(c, ns, ev, evec) bind(C, name="olcao_secular_eqn_pscf")
         import :: c_int, c_ptr
         type(c_ptr), value :: c, ev, evec
         integer(c_int), value :: ns
      end function
      integer(c_int) function c_setup_scf(c) bind(C, name="olcao_setup_scf")
         import :: c_int, c_ptr
         type(c_ptr), value :: c
      end function
      integer(c_int) function c_sec_scf(c, coef, ns, ev, evec) bind(C, name="olcao_secular_eqn_scf")
         import :: c_int, c_ptr
         type(c_ptr), value :: c, coef, ev, evec
         integer(c_int), value :: ns
      end function
      integer(c_int) function c_vale_charge(c, ns, occ, ee, tr, prho) bind(C, name="olcao_vale_charge")
         import :: c_int, c_ptr
         type(c_ptr), value :: c, occ, ee, tr, prho
         integer(c_int), value :: ns
      end function
      integer(c_int) function c_zhegv(n, nb, A, B, w) bind(C, name="olcao_solve_zhegv")
         import :: c_int, c_ptr
         integer(c_int), value :: n, nb
         type(c_ptr), value :: A, B, w
      end function
      integer(c_int) function c_dsygv(n, nb, A, B, w) bind(C, name="olcao_solve_dsygv")
         import :: c_int, c_ptr
         integer(c_int), value :: n, nb
         type(c_ptr), value :: A, B, w
      end function
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
   end interface

contains

   ! Write the library's message to the log unit and stop, as the
   ! reference does on the same conditions.
   subroutine gpuStop(code, where)
      integer(c_int), intent(in) :: code
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      type(c_ptr) :: p
      integer :: n
      p = c_last_error(ctx)
      n = int(c_strlen(p))
      call c_f_pointer(p, msg, [max(n, 1)])
      write (20, *) where, ' failed, status ', code, ': ', msg(1:n)
      stop 'GPU integral/eigen path failed'
   end subroutine gpuStop

   ! Create the device context from the system as parseInput leaves it.
   ! Arrays are 1-based Fortran arrays; atomType/potSiteType hold 1-based
   ! type numbers.
   subroutine gpuInit(atomPos, atomType, typeAlphas, typeNumAlphas, typeNumOrbAlphas, &
         & typeElement, typeNumValeRadial, typeNumCoreRadial, typeValeL, typeCoreL, &
         & typeValeCoef, typeCoreCoef, potAlphas, potNumAlphas, potZ, potNucAlpha, &
         & potElement, potPos, potSiteType, realVectors, kPointsFract, basisCutoff, &
         & gammaBuild, rel, spin)
      real(double), intent(in), target :: atomPos(:,:)          ! (3, nAtoms)
      integer, intent(in) :: atomType(:)
      real(double), intent(in), target :: typeAlphas(:,:)       ! (maxAlphas, nTypes)
      integer, intent(in), target :: typeNumAlphas(:), typeNumOrbAlphas(:,:) ! (4, nTypes)
      integer, intent(in), target :: typeElement(:), typeNumValeRadial(:), typeNumCoreRadial(:)
      integer, intent(in), target :: typeValeL(:,:), typeCoreL(:,:) ! (maxRadial, nTypes)
      real(double), intent(in), target :: typeValeCoef(:,:,:), typeCoreCoef(:,:,:) ! (maxAlphas, maxRadial, nTypes)
      real(double), intent(in), target :: potAlphas(:,:)        ! (maxPotAlphas, nPotTypes)
      integer, intent(in), target :: potNumAlphas(:), potElement(:)
      real(double), intent(in), target :: potZ(:), potNucAlpha(:), potPos(:,:)
      integer, intent(in) :: potSiteType(:)
      real(double), intent(in) :: realVectors(3,3)              ! column i = a_i
      real(double), intent(in), target :: kPointsFract(:,:)     ! (3, nk)
      real(double), intent(in) :: basisCutoff
      integer, intent(in) :: gammaBuild
      integer, intent(in), optional :: rel    ! O_Potential rel (default 0)
      integer, intent(in), optional :: spin   ! O_Potential spin (default 1)
      integer(c_int), allocatable, target :: at0(:), pt0(:)
      type(olcao_system_t) :: s
      type(olcao_options_t) :: o
      integer(c_int) :: st

      if (c_associated(ctx)) call gpuFinalize()
      allocate(at0(size(atomType)), pt0(size(potSiteType)))
      at0 = atomType - 1
      pt0 = potSiteType - 1
      s%nAtoms = size(atomType)
      s%atomPos = c_loc(atomPos)
      s%atomType = c_loc(at0)
      s%nTypes = size(typeNumAlphas)
      s%maxAlphas = size(typeAlphas, 1)
      s%maxRadial = size(typeValeL, 1)
      s%typeNumAlphas = c_loc(typeNumAlphas)
      s%typeAlphas = c_loc(typeAlphas)
      s%typeNumOrbAlphas = c_loc(typeNumOrbAlphas)
      s%typeElement = c_loc(typeElement)
      s%typeNumValeRadial = c_loc(typeNumValeRadial)
      s%typeNumCoreRadial = c_loc(typeNumCoreRadial)
      s%typeValeL = c_loc(typeValeL)
      s%typeCoreL = c_loc(typeCoreL)
      s%typeValeCoef = c_loc(typeValeCoef)
      s%typeCoreCoef = c_loc(typeCoreCoef)
      s%nPotTypes = size(potNumAlphas)
      s%maxPotAlphas = size(potAlphas, 1)
      s%potNumAlphas = c_loc(potNumAlphas)
      s%potAlphas = c_loc(potAlphas)
      s%potZ = c_loc(potZ)
      s%potNucAlpha = c_loc(potNucAlpha)
      s%potElement = c_loc(potElement)
      s%nPotSites = size(potSiteType)
      s%potPos = c_loc(potPos)
      s%potSiteType = c_loc(pt0)
      s%cellVectors = reshape(realVectors, [9])
      s%nk = size(kPointsFract, 2)
      s%kpoints = c_loc(kPointsFract)
      s%basisCutoff = basisCutoff
      s%gamma = gammaBuild
      s%rel = 0
      s%spin = 1
      if (present(rel)) s%rel = rel
      if (present(spin)) s%spin = spin
      o%precision = 0
      o%eigensolver = 0
      o%bloch = 0
      o%graph = 0
      o%chunkItems = 0
      o%segPairs = 0
      o%reserved = 0
      st = c_ctx_create(s, o, ctx)
      if (st /= 0) call gpuStop(st, 'gpuInit')
   end subroutine gpuInit

   subroutine gpuFinalize()
      if (c_associated(ctx)) call c_ctx_destroy(ctx)
      ctx = c_null_ptr
   end subroutine gpuFinalize

   ! valeDim, coreDim, numKPoints, potDim, numCellsReal, pair-cells,
   ! primitive pairs, atom pairs
   subroutine gpuDims(dims)
      integer(c_long_long), intent(out) :: dims(8)
      integer(c_int) :: st
      st = c_get_dims(ctx, dims)
      if (st /= 0) call gpuStop(st, 'gpuDims')
   end subroutine gpuDims

   ! integrals.F90:151
   subroutine gaussOverlapOL(valeVale, coreVale, coreCore)
      complex(double), intent(out), target :: valeVale(:,:,:), coreVale(:,:,:), coreCore(:,:,:)
      integer(c_int) :: st
      st = c_overlap(ctx, 0_c_int, c_loc(valeVale), opt_loc(coreVale), opt_loc(coreCore))
      if (st /= 0) call gpuStop(st, 'gaussOverlapOL')
   end subroutine gaussOverlapOL

   ! integrals.F90:656
   subroutine gaussOverlapKE(valeVale, coreVale, coreCore)
      complex(double), intent(out), target :: valeVale(:,:,:), coreVale(:,:,:), coreCore(:,:,:)
      integer(c_int) :: st
      st = c_overlap(ctx, 1_c_int, c_loc(valeVale), opt_loc(coreVale), opt_loc(coreCore))
      if (st /= 0) call gpuStop(st, 'gaussOverlapKE')
   end subroutine gaussOverlapKE

   ! integrals.F90:1091 (rel = 1; the negated mass-velocity matrices)
   subroutine gaussOverlapMV(valeVale, coreVale, coreCore)
      complex(double), intent(out), target :: valeVale(:,:,:), coreVale(:,:,:), coreCore(:,:,:)
      integer(c_int) :: st
      st = c_overlap(ctx, 2_c_int, c_loc(valeVale), opt_loc(coreVale), opt_loc(coreCore))
      if (st /= 0) call gpuStop(st, 'gaussOverlapMV')
   end subroutine gaussOverlapMV

   ! integrals.F90:1531
   subroutine gaussOverlapNP(valeVale, coreVale, coreCore)
      complex(double), intent(out), target :: valeVale(:,:,:), coreVale(:,:,:), coreCore(:,:,:)
      integer(c_int) :: st
      st = c_np(ctx, c_loc(valeVale), opt_loc(coreVale), opt_loc(coreCore))
      if (st /= 0) call gpuStop(st, 'gaussOverlapNP')
   end subroutine gaussOverlapNP

   ! integrals.F90:2357 (one potential term per call; term is 1-based)
   subroutine elecPotGaussOverlap(term, valeVale, coreVale, coreCore)
      integer, intent(in) :: term
      complex(double), intent(out), target :: valeVale(:,:,:), coreVale(:,:,:), coreCore(:,:,:)
      integer(c_int) :: st
      st = c_ep(ctx, int(term - 1, c_int), c_loc(valeVale), opt_loc(coreVale), opt_loc(coreCore))
      if (st /= 0) call gpuStop(st, 'elecPotGaussOverlap')
   end subroutine elecPotGaussOverlap

   ! integrals.F90:2860; potCoeffs(potDim, spin); the last dimension of the
   ! matrices runs over k within spin (numKPoints * spin)
   subroutine gaussOverlapHamPSCF(potCoeffs, valeVale, coreVale, coreCore)
      real(double), intent(in), target :: potCoeffs(:,:)
      complex(double), intent(out), target :: valeVale(:,:,:), coreVale(:,:,:), coreCore(:,:,:)
      integer(c_int) :: st
      st = c_ham(ctx, c_loc(potCoeffs), c_loc(valeVale), opt_loc(coreVale), opt_loc(coreCore))
      if (st /= 0) call gpuStop(st, 'gaussOverlapHamPSCF')
   end subroutine gaussOverlapHamPSCF

   ! secularEqn.F90:331; energyEigenValues(numStates, numKPoints, spin)
   subroutine secularEqnPSCF(numStates, energyEigenValues)
      integer, intent(in) :: numStates
      real(double), intent(out), target :: energyEigenValues(:,:,:)
      integer(c_int) :: st
      st = c_sec_pscf(ctx, int(numStates, c_int), c_loc(energyEigenValues), c_null_ptr)
      if (st /= 0) call gpuStop(st, 'secularEqnPSCF')
   end subroutine secularEqnPSCF

   ! OLCAO.F90:248-279 (S, KE, NP and every V_j kept on the device)
   subroutine setupSCF()
      integer(c_int) :: st
      st = c_setup_scf(ctx)
      if (st /= 0) call gpuStop(st, 'setupSCF')
   end subroutine setupSCF

   ! secularEqn.F90:37
   subroutine secularEqnSCF(potCoeffs, numStates, energyEigenValues)
      real(double), intent(in), target :: potCoeffs(:,:)       ! (potDim, spin)
      integer, intent(in) :: numStates
      real(double), intent(out), target :: energyEigenValues(:,:,:) ! (numStates, nk, spin)
      integer(c_int) :: st
      st = c_sec_scf(ctx, c_loc(potCoeffs), int(numStates, c_int), &
            & c_loc(energyEigenValues), c_null_ptr)
      if (st /= 0) call gpuStop(st, 'secularEqnSCF')
   end subroutine secularEqnSCF

   ! valeCharge.F90:34 (makeValenceRho, inSCF = 1) after setupSCF and a
   ! solve: electronPopulation (numStates*spin*nk, its order), then
   ! electronEnergy(spin), the traces(4, spin) = charge, nuclear, kinetic,
   ! mass velocity, and potRho(potDim, spin); spin index 1 = total,
   ! 2 = up - down.
   subroutine makeValenceRho(numStates, electronPopulation, electronEnergy, traces, potRho)
      integer, intent(in) :: numStates
      real(double), intent(in), target :: electronPopulation(:)
      real(double), intent(out), target :: electronEnergy(:)
      real(double), intent(out), target :: traces(:,:)
      real(double), intent(out), target :: potRho(:,:)
      integer(c_int) :: st
      st = c_vale_charge(ctx, int(numStates, c_int), c_loc(electronPopulation), &
            & c_loc(electronEnergy), c_loc(traces), c_loc(potRho))
      if (st /= 0) call gpuStop(st, 'makeValenceRho')
   end subroutine makeValenceRho

   ! interfaces.F90:73-120, same dummy arguments: A (upper) in, eigenvectors
   ! out; B (upper) in; the first resultDim eigenvalues returned.
   subroutine solveZHEGV(N, resultDim, A, B, eigenValues)
      integer, intent(in) :: N, resultDim
      complex(double), dimension(N,N), intent(inout), target :: A
      complex(double), dimension(N,N), intent(in) :: B
      real(double), dimension(resultDim), intent(out) :: eigenValues
      complex(double), allocatable, target :: Bw(:,:)
      real(double), allocatable, target :: w(:)
      integer(c_int) :: st
      allocate(Bw(N,N), w(N))
      Bw = B
      st = c_zhegv(int(N, c_int), 1_c_int, c_loc(A), c_loc(Bw), c_loc(w))
      if (st /= 0) then
         write (20,*) 'ZHEGV error code=', st
         stop
      endif
      eigenValues(:resultDim) = w(:resultDim)
   end subroutine solveZHEGV

   ! interfaces.F90:152-200
   subroutine solveDSYGV(valeDim, numStates, valeValeGamma, valeValeOLGamma, eigenValues)
      integer, intent(in) :: valeDim, numStates
      real(double), dimension(valeDim,valeDim), intent(inout), target :: valeValeGamma
      real(double), dimension(valeDim,valeDim), intent(in) :: valeValeOLGamma
      real(double), dimension(valeDim), intent(out) :: eigenValues
      real(double), allocatable, target :: Bw(:,:), w(:)
      integer(c_int) :: st
      allocate(Bw(valeDim,valeDim), w(valeDim))
      Bw = valeValeOLGamma
      st = c_dsygv(int(valeDim, c_int), 1_c_int, c_loc(valeValeGamma), c_loc(Bw), c_loc(w))
      if (st /= 0) then
         write (20,*) 'DSYGV error code=', st
         stop
      endif
      eigenValues(:numStates) = w(:numStates)
   end subroutine solveDSYGV

   function opt_loc(a) result(p)
      complex(double), intent(in), target :: a(:,:,:)
      type(c_ptr) :: p
      if (size(a) > 0) then
         p = c_loc(a)
      else
         p = c_null_ptr
      endif
   end function opt_loc
end module O_GPUFacade

! C-callable entries of the facade, used by the Python harness to drive the
! Fortran call path itself (ctypes -> Fortran -> C-ABI -> GPU).
module O_GPUFacadeCheck
   use iso_c_binding
   use O_GPUFacade
   implicit none
contains
   subroutine facade_solve_zhegv(n, A, B, w) bind(C, name="olcao_facade_solve_zhegv")
      integer(c_int), value :: n
      complex(c_double_complex), intent(inout) :: A(n,n)
      complex(c_double_complex), intent(in) :: B(n,n)
      real(c_double), intent(out) :: w(n)
      call solveZHEGV(int(n), int(n), A, B, w)
   end subroutine facade_solve_zhegv

   subroutine facade_solve_dsygv(n, A, B, w) bind(C, name="olcao_facade_solve_dsygv")
      integer(c_int), value :: n
      real(c_double), intent(inout) :: A(n,n)
      real(c_double), intent(in) :: B(n,n)
      real(c_double), intent(out) :: w(n)
      call solveDSYGV(int(n), int(n), A, B, w)
   end subroutine facade_solve_dsygv
end module O_GPUFacadeCheck
