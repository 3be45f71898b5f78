! iso_c_binding interface to libtonto_gpu (include/tonto_gpu.h) for Tonto's Fortran.
! A Tonto build links libtonto_gpu.so and calls these from MOLECULE.FOCK:make_r_JK_engine
! (foofiles/molecule.fock.foo:1365-1505); see INTEGRATION.md for the shim.
module tonto_gpu_mod
   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: TGPU_OK = 0
   integer(c_int), parameter :: TGPU_JK_MIXED = 1

   type, bind(c) :: tgpu_system
      integer(c_int) :: n_atom
      integer(c_int) :: n_shell
      type(c_ptr)    :: atom_xyz
      type(c_ptr)    :: atom_Z
      type(c_ptr)    :: shell_l
      type(c_ptr)    :: shell_atom
      type(c_ptr)    :: shell_nprim
      type(c_ptr)    :: exps
      type(c_ptr)    :: coefs
   end type

   interface
      integer(c_int) function tgpu_create(sys, ctx) bind(c, name="tgpu_create")
         import :: c_int, c_ptr, tgpu_system
         type(tgpu_system), intent(in) :: sys
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function tgpu_destroy(ctx) bind(c, name="tgpu_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function tgpu_last_error(ctx) bind(c, name="tgpu_last_error")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function tgpu_set_cutoffs(ctx, schwarz, jp, kp, pair) bind(c, name="tgpu_set_cutoffs")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: schwarz, jp, kp, pair
      end function
      integer(c_int) function tgpu_schwarz(ctx, max_I) bind(c, name="tgpu_schwarz")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: max_I(*)
      end function
      integer(c_int) function tgpu_make_r_JK(ctx, P, J, K, flags) bind(c, name="tgpu_make_r_JK")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in)  :: P(*)
         real(c_double), intent(out) :: J(*), K(*)
         integer(c_int), value :: flags
      end function
      integer(c_int) function tgpu_make_u_JK(ctx, Pa, Pb, Ja, Jb, Ka, Kb, flags) bind(c, name="tgpu_make_u_JK")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in)  :: Pa(*), Pb(*)
         real(c_double), intent(out) :: Ja(*), Jb(*), Ka(*), Kb(*)
         integer(c_int), value :: flags
      end function
   end interface

   interface
      integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
   end interface

contains

   ! The text of tgpu_last_error(ctx) as a Fortran string.
   function tgpu_error_message(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = tgpu_last_error(ctx)
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function

   ! Tonto's DIE for a failed call: stop with the library's message.
   subroutine tgpu_die(ctx, rc, what)
      type(c_ptr), intent(in) :: ctx
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      write(*, '(a, a, i0, a, a)') what, ": libtonto_gpu error ", rc, ": ", tgpu_error_message(ctx)
      error stop 1
   end subroutine

   ! Replacement body for make_r_JK_engine(J,K,P): J and K come back after Tonto's
   ! post-processing, so make_r_Fock_mx's F = J - K/2 (+ core) is unchanged.
   subroutine gpu_make_r_JK_engine(ctx, J, K, P)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(in)  :: P(:,:)
      real(c_double), intent(out) :: J(:,:), K(:,:)
      integer(c_int) :: rc
      rc = tgpu_make_r_JK(ctx, P, J, K, 0_c_int)
      if (rc /= TGPU_OK) call tgpu_die(ctx, rc, "make_r_JK_engine")
   end subroutine

   ! Replacement body for make_u_JK_engine(J,K,P) (molecule.fock.foo:2558-2712):
   ! J.a, J.b, K.a, K.b as Tonto's engine leaves them.
   subroutine gpu_make_u_JK_engine(ctx, Pa, Pb, Ja, Jb, Ka, Kb)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(in)  :: Pa(:,:), Pb(:,:)
      real(c_double), intent(out) :: Ja(:,:), Jb(:,:), Ka(:,:), Kb(:,:)
      integer(c_int) :: rc
      rc = tgpu_make_u_JK(ctx, Pa, Pb, Ja, Jb, Ka, Kb, 0_c_int)
      if (rc /= TGPU_OK) call tgpu_die(ctx, rc, "make_u_JK_engine")
   end subroutine

end module
