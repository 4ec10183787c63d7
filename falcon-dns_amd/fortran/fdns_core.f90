! fdns_core.f90 -- ISO_C_BINDING interface of libfdns.so (include/fdns.h).
!
! The Fortran side of the drop-in boundary: a Fortran driver keeps the
! reference's namelists and output files and hands the time stepping
! (solver.F90:68-88, timeloop_imex / timeloop_imexrk) to the MI355X library
! through these interfaces.  Spectral fields are the reference's
! complex(dp) (Nm_max+1, Nr_max) arrays, passed by address.
module fdns_core
   use iso_c_binding
   implicit none
   private

   integer, parameter, public :: FDNS_OK = 0
   integer, parameter, public :: FDNS_SOLVE_AUTO = 0, FDNS_SOLVE_LU = 1, FDNS_SOLVE_INVERSE = 2

   ! fdns_params of include/fdns.h, field for field
   type, bind(C), public :: fdns_params
      integer(c_int32_t) :: Nr
      integer(c_int32_t) :: Nm
      real(c_double) :: eta
      real(c_double) :: Ra
      real(c_double) :: Pr
      real(c_double) :: dt
      character(kind=c_char) :: mBC(8)
      character(kind=c_char) :: buo_tscheme(8)
      character(kind=c_char) :: time_scheme_type(16)
      character(kind=c_char) :: time_scheme_imp(16)
      character(kind=c_char) :: time_scheme_exp(16)
      integer(c_int32_t) :: l_restart
      integer(c_int32_t) :: n_KE
      integer(c_int32_t) :: solver
      integer(c_int32_t) :: l_vartimestep
      integer(c_int32_t) :: l_optimizedt
      real(c_double) :: CFL
      real(c_double) :: dt_coef
      real(c_double) :: dt_max
      integer(c_int32_t) :: lagpts
   end type fdns_params

   public :: set_cstr, fdns_errmsg

   interface
      integer(c_int) function fdns_create(p, run) bind(C, name="fdns_create")
         import :: c_int, c_ptr, fdns_params
         type(fdns_params), intent(in) :: p
         type(c_ptr), intent(out) :: run
      end function fdns_create

      integer(c_int) function fdns_destroy(run) bind(C, name="fdns_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: run
      end function fdns_destroy

      integer(c_int) function fdns_set_state(run, tFR, omgFR, urFR, upFR) bind(C, name="fdns_set_state")
         import :: c_int, c_ptr
         type(c_ptr), value :: run, tFR, omgFR, urFR, upFR
      end function fdns_set_state

      integer(c_int) function fdns_get_state(run, tFR, omgFR, urFR, upFR, psi) bind(C, name="fdns_get_state")
         import :: c_int, c_ptr
         type(c_ptr), value :: run, tFR, omgFR, urFR, upFR, psi
      end function fdns_get_state

      integer(c_int) function fdns_seed_state(run, seed, amp) bind(C, name="fdns_seed_state")
         import :: c_int, c_ptr, c_int64_t, c_double
         type(c_ptr), value :: run
         integer(c_int64_t), value :: seed
         real(c_double), value :: amp
      end function fdns_seed_state

      integer(c_int) function fdns_init_perturbation(run, n_init, ampT) bind(C, name="fdns_init_perturbation")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: run
         integer(c_int), value :: n_init
         real(c_double), value :: ampT
      end function fdns_init_perturbation

      integer(c_int) function fdns_restart_dt(run, n, dt_array) bind(C, name="fdns_restart_dt")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: run
         integer(c_int), value :: n
         real(c_double), intent(in) :: dt_array(*)
      end function fdns_restart_dt

      integer(c_int) function fdns_step(run, nsteps) bind(C, name="fdns_step")
         import :: c_int, c_ptr
         type(c_ptr), value :: run
         integer(c_int), value :: nsteps
      end function fdns_step

      integer(c_int) function fdns_get_dt(run, dt_new, dt_array) bind(C, name="fdns_get_dt")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: run
         real(c_double), intent(out) :: dt_new
         type(c_ptr), value :: dt_array
      end function fdns_get_dt

      integer(c_int) function fdns_diag(run, out6) bind(C, name="fdns_diag")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: run
         real(c_double), intent(out) :: out6(6)
      end function fdns_diag

      integer(c_int) function fdns_budget(run, out4, upbal) bind(C, name="fdns_budget")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: run
         real(c_double), intent(out) :: out4(4)
         type(c_ptr), value :: upbal
      end function fdns_budget

      integer(c_int) function fdns_refine_state(Nm_old, Nr_old, tF, oF, rF, pF, Nm_new, Nr_new, tN, oN, rN, pN) &
            bind(C, name="fdns_refine_state")
         import :: c_int, c_ptr
         integer(c_int), value :: Nm_old, Nr_old, Nm_new, Nr_new
         type(c_ptr), value :: tF, oF, rF, pF, tN, oN, rN, pN
      end function fdns_refine_state

      integer(c_int) function fdns_spectra(run, KEspec, Tspec) bind(C, name="fdns_spectra")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: run
         real(c_double), intent(out) :: KEspec(*), Tspec(*)
      end function fdns_spectra

      function fdns_strerror(code) bind(C, name="fdns_strerror")
         import :: c_int, c_ptr
         integer(c_int), value :: code
         type(c_ptr) :: fdns_strerror
      end function fdns_strerror

      integer(c_size_t) function fdns_job_footprint(p) bind(C, name="fdns_job_footprint")
         import :: c_size_t, fdns_params
         type(fdns_params), intent(in) :: p
      end function fdns_job_footprint

      integer(c_int) function fdns_scheme_orders(p, n_imp, n_exp) bind(C, name="fdns_scheme_orders")
         import :: c_int, fdns_params
         type(fdns_params), intent(in) :: p
         integer(c_int), intent(out) :: n_imp, n_exp
      end function fdns_scheme_orders

      integer(c_int) function fdns_restart(run, nstates, tFR, omgFR, urFR, upFR, time) bind(C, name="fdns_restart")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: run, tFR, omgFR, urFR, upFR
         integer(c_int), value :: nstates
         real(c_double), value :: time
      end function fdns_restart

      integer(c_int) function fdns_get_history(run, imp_temp, exp_temp, imp_vort, exp_vort, imp_ub, exp_ub) &
         & bind(C, name="fdns_get_history")
         import :: c_int, c_ptr
         type(c_ptr), value :: run, imp_temp, exp_temp, imp_vort, exp_vort, imp_ub, exp_ub
      end function fdns_get_history

      integer(c_int) function fdns_write_checkpoint(run, path) bind(C, name="fdns_write_checkpoint")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: run
         character(kind=c_char), intent(in) :: path(*)
      end function fdns_write_checkpoint

      integer(c_int) function fdns_write_snapshot(run, path) bind(C, name="fdns_write_snapshot")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: run
         character(kind=c_char), intent(in) :: path(*)
      end function fdns_write_snapshot
   end interface

   public :: fdns_create, fdns_destroy, fdns_set_state, fdns_get_state, fdns_seed_state, fdns_step, &
             fdns_init_perturbation, fdns_restart_dt, &
             fdns_diag, fdns_budget, fdns_spectra, fdns_refine_state, fdns_get_dt, fdns_strerror, fdns_job_footprint, fdns_scheme_orders, fdns_restart, &
             fdns_get_history, fdns_write_checkpoint, fdns_write_snapshot

contains

   ! Copy a Fortran string into a NUL-terminated C char array.
   subroutine set_cstr(dst, src)
      character(kind=c_char), intent(out) :: dst(:)
      character(len=*), intent(in) :: src
      integer :: i, n
      dst = c_null_char
      n = min(len_trim(src), size(dst) - 1)
      do i = 1, n
         dst(i) = src(i:i)
      end do
   end subroutine set_cstr

   ! The library's message for a status code, as a Fortran string.
   function fdns_errmsg(code) result(msg)
      integer(c_int), intent(in) :: code
      character(len=256) :: msg
      character(kind=c_char), pointer :: s(:)
      type(c_ptr) :: p
      integer :: i
      msg = ' '
      p = fdns_strerror(code)
      if (.not. c_associated(p)) return
      call c_f_pointer(p, s, [256])
      do i = 1, 256
         if (s(i) == c_null_char) exit
         msg(i:i) = s(i)
      end do
   end function fdns_errmsg

end module fdns_core
