! Fortran host program over the GPU facade (module O_GPUFacade), calling it
! in the order of the reference driver's call sites:
!   setupSCF   src/uolcao/OLCAO.F90:248-279  gaussOverlapOL, KE, NP, every
!                                            potential term, then the device-
!                                            resident term set (setupSCF)
!   mainSCF    OLCAO.F90:482                 secularEqnSCF
!   intgPSCF   OLCAO.F90:645-662             gaussOverlapOL, gaussOverlapHamPSCF
!   bandPSCF   OLCAO.F90:758                 secularEqnPSCF
! and the drop-in solver solveZHEGV (interfaces.F90:73) on the first
! k-point's matrices, as secularEqn.F90:227 calls it.
!
! Test program: reads a system (what parseInput leaves in the module
! globals) from the file named by the first argument, writes eigenvalues and
! the first k-point's S and H to the file named by the second.
program facade_driver
   use O_GPUFacade
   use iso_c_binding, only: c_long_long
   implicit none
   integer, parameter :: double = kind(0.0d0)
   character(len=1024) :: inFile, outFile
   integer :: nAtoms, nTypes, maxAlphas, maxRadial, nPotTypes, maxPotAlphas, nPotSites
   integer :: nk, gammaBuild, rel, spin, potDim, numStates
   real(double) :: basisCutoff, realVectors(3,3)
   real(double), allocatable :: atomPos(:,:), typeAlphas(:,:), typeValeCoef(:,:,:)
   real(double), allocatable :: typeCoreCoef(:,:,:), potAlphas(:,:), potZ(:), potNucAlpha(:)
   real(double), allocatable :: potPos(:,:), kPointsFract(:,:), potCoeffs(:,:)
   integer, allocatable :: atomType(:), typeNumAlphas(:), typeNumOrbAlphas(:,:), typeElement(:)
   integer, allocatable :: typeNumValeRadial(:), typeNumCoreRadial(:), typeValeL(:,:)
   integer, allocatable :: typeCoreL(:,:), potNumAlphas(:), potElement(:), potSiteType(:)
   integer(c_long_long) :: dims(8)
   integer :: n, nc, j, u
   complex(double), allocatable :: S(:,:,:), M(:,:,:), cv(:,:,:), cc(:,:,:), H(:,:,:)
   complex(double), allocatable :: Hk(:,:), Sk(:,:)
   real(double), allocatable :: evA(:,:,:), evB(:,:,:), evZ(:)

   call get_command_argument(1, inFile)
   call get_command_argument(2, outFile)
   open (newunit=u, file=trim(inFile), status='old', action='read')
   read (u, *) nAtoms, nTypes, maxAlphas, maxRadial, nPotTypes, maxPotAlphas, nPotSites
   read (u, *) nk, gammaBuild, rel, spin, potDim, numStates, basisCutoff
   allocate (atomPos(3,nAtoms), atomType(nAtoms), typeAlphas(maxAlphas,nTypes))
   allocate (typeNumAlphas(nTypes), typeNumOrbAlphas(4,nTypes), typeElement(nTypes))
   allocate (typeNumValeRadial(nTypes), typeNumCoreRadial(nTypes))
   allocate (typeValeL(maxRadial,nTypes), typeCoreL(maxRadial,nTypes))
   allocate (typeValeCoef(maxAlphas,maxRadial,nTypes), typeCoreCoef(maxAlphas,maxRadial,nTypes))
   allocate (potAlphas(maxPotAlphas,nPotTypes), potNumAlphas(nPotTypes), potZ(nPotTypes))
   allocate (potNucAlpha(nPotTypes), potElement(nPotTypes), potPos(3,nPotSites))
   allocate (potSiteType(nPotSites), kPointsFract(3,nk), potCoeffs(potDim,spin))
   read (u, *) atomPos
   read (u, *) atomType
   read (u, *) typeAlphas
   read (u, *) typeNumAlphas
   read (u, *) typeNumOrbAlphas
   read (u, *) typeElement
   read (u, *) typeNumValeRadial
   read (u, *) typeNumCoreRadial
   read (u, *) typeValeL
   read (u, *) typeCoreL
   read (u, *) typeValeCoef
   read (u, *) typeCoreCoef
   read (u, *) potAlphas
   read (u, *) potNumAlphas
   read (u, *) potZ
   read (u, *) potNucAlpha
   read (u, *) potElement
   read (u, *) potPos
   read (u, *) potSiteType
   read (u, *) realVectors
   read (u, *) kPointsFract
   read (u, *) potCoeffs
   close (u)

   call gpuInit(atomPos, atomType, typeAlphas, typeNumAlphas, typeNumOrbAlphas, &
         & typeElement, typeNumValeRadial, typeNumCoreRadial, typeValeL, typeCoreL, &
         & typeValeCoef, typeCoreCoef, potAlphas, potNumAlphas, potZ, potNucAlpha, &
         & potElement, potPos, potSiteType, realVectors, kPointsFract, basisCutoff, &
         & gammaBuild, rel, spin)
   call gpuDims(dims)
   n = int(dims(1))
   nc = int(dims(2))
   allocate (S(n,n,nk), M(n,n,nk), H(n,n,nk*spin), cv(nc,n,nk*spin), cc(nc,nc,nk*spin))
   allocate (evA(numStates,nk,spin), evB(numStates,nk,spin), evZ(numStates))

   ! setupSCF (OLCAO.F90:248-279)
   call gaussOverlapOL(S, cv(:,:,1:nk), cc(:,:,1:nk))
   call gaussOverlapKE(M, cv(:,:,1:nk), cc(:,:,1:nk))
   call gaussOverlapNP(M, cv(:,:,1:nk), cc(:,:,1:nk))
   do j = 1, potDim
      call elecPotGaussOverlap(j, M, cv(:,:,1:nk), cc(:,:,1:nk))
   enddo
   call setupSCF()
   ! mainSCF (OLCAO.F90:482)
   call secularEqnSCF(potCoeffs, numStates, evA)
   ! intgPSCF (OLCAO.F90:645-662) and bandPSCF (:758)
   call gaussOverlapOL(S, cv(:,:,1:nk), cc(:,:,1:nk))
   call gaussOverlapHamPSCF(potCoeffs, H, cv, cc)
   call secularEqnPSCF(numStates, evB)
   ! the drop-in solver on the first k-point (secularEqn.F90:227)
   allocate (Hk(n,n), Sk(n,n))
   Hk = H(:,:,1)
   Sk = S(:,:,1)
   call solveZHEGV(n, numStates, Hk, Sk, evZ)

   open (newunit=u, file=trim(outFile), status='replace', action='write')
   write (u, '(3i8)') n, nk, spin
   write (u, '(es26.17e3)') evA
   write (u, '(es26.17e3)') evB
   write (u, '(es26.17e3)') evZ
   write (u, '(es26.17e3)') S(:,:,1)
   write (u, '(es26.17e3)') H(:,:,1)
   close (u)
   call gpuFinalize()
end program facade_driver
