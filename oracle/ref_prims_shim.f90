! C-callable driver around the reference's primitive-integral subroutines
! (src/uolcao/gaussIntegrals.f90).  Test infrastructure only: it is linked
! into oracle/_ref/libolcao_ref_prims.so, which exists solely to pin the
! oracle's primitive integrals and Boys function against the reference's own
! compiled code.  Nothing in the product loads it.
module ref_prims_shim
   use iso_c_binding
   use O_GaussianIntegrals
   implicit none
contains
   ! op: 0 overlap2CIntg, 1 kinetic2CIntg, 2 electron3CIntg,
   !     3 nuclear3CIntg, 4 massvel2CIntg.  sh is column-major (16,16); only
   !     the block selected by sw is defined, the rest is left at zero.
   subroutine ref_prim(op, a1, a2, a3, A, B, C, sw, sh) bind(C, name="ref_prim")
      integer(c_int), value :: op, sw
      real(c_double), value :: a1, a2, a3
      real(c_double), intent(in) :: A(3), B(3), C(3)
      real(c_double), intent(inout) :: sh(16,16)
      sh(:,:) = 0.0d0
      select case (op)
      case (0)
         call overlap2CIntg(a1, a2, A, B, sw, sh)
      case (1)
         call kinetic2CIntg(a1, a2, A, B, sw, sh)
      case (2)
         call electron3CIntg(a1, a2, a3, A, B, C, sw, sh)
      case (3)
         call nuclear3CIntg(a1, a2, a3, A, B, C, sw, sh)
      case (4)
         call massvel2CIntg(a1, a2, A, B, sw, sh)
      end select
   end subroutine ref_prim

   ! Three-component operators: op 5 dipole3CIntg (dipole about C), op 6
   ! momentum2CIntg (gradient on the ket).  sh is column-major (16,16,3).
   subroutine ref_prim3(op, a1, a2, A, B, C, sw, sh) bind(C, name="ref_prim3")
      integer(c_int), value :: op, sw
      real(c_double), value :: a1, a2
      real(c_double), intent(in) :: A(3), B(3), C(3)
      real(c_double), intent(inout) :: sh(16,16,3)
      sh(:,:,:) = 0.0d0
      select case (op)
      case (5)
         call dipole3CIntg(a1, a2, A, B, C, sw, sh)
      case (6)
         call momentum2CIntg(a1, a2, A, B, sw, sh)
      end select
   end subroutine ref_prim3

   ! The force-derivative primitives of computeForceIntg: op 10
   ! dkinetic2CIntg, op 11 dnuclear3CIntgCB, op 12 delectron3CIntgCB (a3 and
   ! C unused by op 10).  sh is column-major (16,16,3).
   subroutine ref_primd(op, a1, a2, a3, A, B, C, sw, sh) bind(C, name="ref_primd")
      integer(c_int), value :: op, sw
      real(c_double), value :: a1, a2, a3
      real(c_double), intent(in) :: A(3), B(3), C(3)
      real(c_double), intent(inout) :: sh(16,16,3)
      sh(:,:,:) = 0.0d0
      select case (op)
      case (10)
         call dkinetic2CIntg(a1, a2, A, B, sw, sh)
      case (11)
         call dnuclear3CIntgCB(a1, a2, a3, A, B, C, sw, sh)
      case (12)
         call delectron3CIntgCB(a1, a2, a3, A, B, C, sw, sh)
      end select
   end subroutine ref_primd

   ! KOverlap2CIntg: the overlap with exp(-i deltaK . r) between the two
   ! Gaussians; the complex (16,16) block split into its real and imaginary
   ! parts (column-major).
   subroutine ref_primk(a1, a2, A, B, dK, sw, shr, shi) bind(C, name="ref_primk")
      integer(c_int), value :: sw
      real(c_double), value :: a1, a2
      real(c_double), intent(in) :: A(3), B(3), dK(3)
      real(c_double), intent(inout) :: shr(16,16), shi(16,16)
      complex(kind=8) :: sh(16,16)
      sh(:,:) = (0.0d0, 0.0d0)
      call KOverlap2CIntg(a1, a2, A, B, dK, sw, sh)
      shr(:,:) = real(sh(:,:), 8)
      shi(:,:) = aimag(sh(:,:))
   end subroutine ref_primk

   subroutine ref_boys(T, F) bind(C, name="ref_boys")
      real(c_double), value :: T
      real(c_double), intent(out) :: F(7)
      call boys(T, F)
   end subroutine ref_boys
end module ref_prims_shim
