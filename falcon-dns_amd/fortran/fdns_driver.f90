! fdns_driver.f90 -- Fortran driver of the MI355X time-stepping core.
!
! Reads the reference's input file (first command-line argument) with the
! same namelist groups, variables and defaults (namelists.F90:51-140), builds
! the seeded initial state, advances n_time_steps steps on the GPU through
! libfdns.so and appends the kinetic-energy and Nusselt series every n_KE
! steps to KE_<tag> and Nusselt_<tag> in the reference's formats
! (output.F90:30-40, 252-263).
!
! One extra namelist group, &fdns, carries what the reference does not have:
!   seed            -- seed of the build-defined random initial perturbation
!                      (amplitude ampT on modes 1..Nm_max)
!   l_const_history -- start the multistep schemes as a restart from identical
!                      checkpoints of the initial state (constant histories)
!                      instead of the zero histories of a fresh start.
!   solver          -- "auto" (default), "lu" (dgetrs order) or "inverse"
!                      (resident operator inverses), see include/fdns.h.
! Not handled here (the driver stops with a message): l_restart (reading the
! reference's checkpoint files), l_vartimestep (CFL-controlled dt).
program fdns_driver
   use iso_c_binding
   use fdns_core
   implicit none

   integer, parameter :: dp = kind(1.0d0)
   ! &grid
   integer :: Nr_max, Nm_max
   ! &physics
   real(dp) :: eta, Ra, Pr, ampT
   character(len=100) :: mBC, buo_tscheme
   logical :: l_add_pert
   integer :: lagpts, n_init
   ! &timecontrol
   integer :: n_time_steps, n_restart, n_restart_point, n_snapshot_point
   real(dp) :: dt, totaltime, dt_coef, dt_max, CFL
   logical :: l_vartimestep, l_imexrk_started, l_restart, l_optimizedt
   character(len=100) :: time_scheme_type, time_scheme_imp, time_scheme_exp
   ! &output
   character(len=100) :: tag
   integer :: n_checkpoint, n_snapshot, n_KE, n_KEspec
   ! &fdns
   integer(c_int64_t) :: seed
   logical :: l_const_history
   character(len=16) :: solver

   namelist /grid/ Nr_max, Nm_max
   namelist /physics/ eta, Ra, Pr, mBC, ampT, l_add_pert, lagpts, n_init, buo_tscheme
   namelist /timecontrol/ n_time_steps, dt, totaltime, l_vartimestep, time_scheme_type, time_scheme_imp, &
      & time_scheme_exp, l_imexrk_started, dt_coef, dt_max, CFL, l_restart, l_optimizedt, n_restart, &
      & n_restart_point, n_snapshot_point
   namelist /output/ tag, n_checkpoint, n_snapshot, n_KE, n_KEspec
   namelist /fdns/ seed, l_const_history, solver

   type(fdns_params) :: p
   type(c_ptr) :: run
   integer(c_int) :: rc
   real(c_double) :: d(6)
   character(len=512) :: fname
   integer :: u, res, done, chunk, ke_unit, nu_unit
   integer(8) :: t0, t1, rate

   ! defaults of defaultNamelists (namelists.F90:117-140)
   Nr_max = 20; Nm_max = 20
   eta = 0.35_dp; Ra = 1000.0_dp; Pr = 1.0_dp; mBC = "SF"; ampT = 0.00001_dp
   l_add_pert = .false.; lagpts = 14; buo_tscheme = "IMP"; n_init = 0
   n_time_steps = 100; dt = 0.01_dp; totaltime = 100.0_dp; l_vartimestep = .false.; l_optimizedt = .false.
   time_scheme_type = "IMEXRK"; time_scheme_imp = "ARS222"; time_scheme_exp = "expARS222"
   l_imexrk_started = .true.; dt_coef = 2.0_dp; dt_max = 3.e-5_dp; CFL = 0.9_dp
   l_restart = .false.; n_restart = 0; n_restart_point = 0; n_snapshot_point = 0
   tag = "test"; n_checkpoint = 1000; n_snapshot = 1000; n_KE = 1; n_KEspec = 1000
   seed = 0; l_const_history = .false.; solver = "auto"

   if (command_argument_count() < 1) then
      write (0, '(a)') "usage: fdns_driver <input namelist file>"
      stop 2
   end if
   call get_command_argument(1, fname)
   open (newunit=u, file=trim(fname), status="old", action="read", iostat=res)
   if (res /= 0) then
      write (0, '(a)') "input namelist file not found: "//trim(fname)
      stop 2
   end if
   read (u, nml=grid, iostat=res); rewind (u)
   read (u, nml=physics, iostat=res); rewind (u)
   read (u, nml=timecontrol, iostat=res); rewind (u)
   read (u, nml=output, iostat=res); rewind (u)
   read (u, nml=fdns, iostat=res)
   close (u)

   if (l_restart) then
      write (0, '(a)') "l_restart=.TRUE. (reading checkpoint files) is not supported by this driver"
      stop 2
   end if
   if (l_vartimestep) then
      write (0, '(a)') "l_vartimestep=.TRUE. is not supported: the GPU core steps with a constant dt"
      stop 2
   end if

   p%Nr = Nr_max
   p%Nm = Nm_max
   p%eta = eta
   p%Ra = Ra
   p%Pr = Pr
   p%dt = dt
   call set_cstr(p%mBC, mBC)
   call set_cstr(p%buo_tscheme, buo_tscheme)
   call set_cstr(p%time_scheme_type, time_scheme_type)
   call set_cstr(p%time_scheme_imp, time_scheme_imp)
   call set_cstr(p%time_scheme_exp, time_scheme_exp)
   p%l_restart = merge(1, 0, l_const_history)
   p%n_KE = 0
   select case (trim(solver))
   case ("auto")
      p%solver = FDNS_SOLVE_AUTO
   case ("lu")
      p%solver = FDNS_SOLVE_LU
   case ("inverse")
      p%solver = FDNS_SOLVE_INVERSE
   case default
      write (0, '(a)') "unknown solver: "//trim(solver)
      stop 2
   end select

   rc = fdns_create(p, run)
   if (rc /= FDNS_OK) call fail("fdns_create", rc)
   rc = fdns_seed_state(run, seed, ampT)
   if (rc /= FDNS_OK) call fail("fdns_seed_state", rc)

   open (newunit=ke_unit, file="KE_"//trim(tag), status="unknown", position="append", action="write")
   open (newunit=nu_unit, file="Nusselt_"//trim(tag), status="unknown", position="append", action="write")

   call system_clock(t0, rate)
   done = 0
   do while (done < n_time_steps)
      chunk = min(max(n_KE, 1), n_time_steps - done)
      rc = fdns_step(run, chunk)
      if (rc /= FDNS_OK) call fail("fdns_step", rc)
      done = done + chunk
      rc = fdns_diag(run, d)
      if (rc /= FDNS_OK) call fail("fdns_diag", rc)
      ! KE_<tag>: dt, time, KE, KE of u_r, KE of u_phi (output.F90:252)
      write (ke_unit, '(5es20.12)') dt, d(1), d(2), d(5), d(6)
      ! Nusselt_<tag>: time, Nu_t, Nu_b (output.F90:261)
      write (nu_unit, '(3ES20.12)') d(1), d(3), d(4)
      if (d(2) > 1.0e10_dp) then
         write (0, '(a)') "kinetic energy above 1e10: stopping (output.F90:321-323)"
         exit
      end if
      if (d(1) > totaltime) exit
   end do
   call system_clock(t1)

   close (ke_unit)
   close (nu_unit)
   write (6, '(a,i0,a,es12.5,a)') "steps: ", done, "  wall time [s]: ", real(t1 - t0, dp) / real(rate, dp), ""
   rc = fdns_destroy(run)

contains

   subroutine fail(what, code)
      character(len=*), intent(in) :: what
      integer(c_int), intent(in) :: code
      write (0, '(a,a,i0,a,a)') what, " failed (", code, "): ", trim(fdns_errmsg(code))
      stop 1
   end subroutine fail

end program fdns_driver
