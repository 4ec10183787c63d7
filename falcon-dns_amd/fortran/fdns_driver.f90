! fdns_driver.f90 -- Fortran driver of the MI355X time-stepping core.
!
! Reads the reference's input file (first command-line argument) with the
! same namelist groups, variables and defaults (namelists.F90:51-146), builds
! the initial state (seeded, or read from checkpoint files when l_restart),
! advances steps n_restart+1 .. n_restart+n_time_steps on the GPU through
! libfdns.so, and writes what the reference's time loop writes at the same
! step numbers (steptime.F90:253-292, steptime_imexrk.F90:191-228):
!   snapshot_plot_NNNNN  every n_snapshot steps   (store_snapshot, output.F90:451-516)
!   checkpoint_NNNNN     every n_checkpoint steps (store_checkpoint, output.F90:359-450)
!   KE_<tag>, Nusselt_<tag>, power_<tag>, upbalance_<tag>  every n_KE steps
!                        (writeKE_spectral, output.F90:30-40, 122-324)
!   KEspec_<tag>, Temp_spec_<tag>  every n_KEspec steps (calculate_spectra,
!                        output.F90:327-355)
! Checkpoints and snapshots are written here with Fortran unformatted I/O, as
! the reference writes them; the numbering continues from n_restart_point and
! n_snapshot_point.  Output stops once the time reaches totaltime.
!
! l_restart = .TRUE. reads checkpoint_NNNNN files like get_checkpoint_data
! (init.F90:447-592, 865-935): with l_imexrk_started an IMEX scheme reads the
! n_order_tscheme_imp consecutive checkpoints ending at n_restart_point and
! rebuilds its histories from their states (steptime.F90:87-193); otherwise
! one checkpoint.  l_add_pert adds ampT to the restarted temperature
! (init.F90:188-190).  A single checkpoint of a smaller Nm_max / Nr_max is
! carried to the namelist's grid as init.F90:722-861 does
! (fdns_refine_state).
!
! A fresh start (l_restart = .FALSE.) builds the reference's initial state
! (init_fields + init_perturbation, init.F90:164-399): the conductive profile
! plus the perturbation n_init selects, of amplitude ampT.
!
! One extra namelist group, &fdns, carries what the reference does not have:
!   seed            -- when given (>= 0), a fresh start takes the build-defined
!                      random initial perturbation of that seed (amplitude
!                      ampT on modes 1..Nm_max) instead of n_init's
!   l_const_history -- start the multistep schemes as a restart from identical
!                      checkpoints of the initial state (constant histories)
!                      instead of the zero histories of a fresh start.
!   solver          -- "auto" (default), "lu" (dgetrs order) or "inverse"
!                      (resident operator inverses), see include/fdns.h.
! l_vartimestep = .TRUE. steps with the CFL-limited dt of compute_new_dt
! (steptime.F90:313-370); the KE file's first column is each row's dt.
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
   integer(c_int) :: rc, n_imp, n_exp
   real(c_double) :: d(6), dt_new, bud(4)
   real(c_double), allocatable, target :: upbal(:,:), kesp(:), tsp(:)
   character(len=512) :: fname
   integer :: u, res, n_step, n_end, nxt, ke_unit, nu_unit, count_chkpnt, count_snap
   integer :: pw_unit, up_unit, ks_unit, ts_unit, j
   integer(8) :: t0, t1, rate
   ! restart: the checkpoint states, oldest first, the newest one's time and
   ! the dt_array get_checkpoint_data builds from them
   complex(dp), allocatable, target :: tS(:,:,:), oS(:,:,:), rS(:,:,:), pS(:,:,:)
   real(dp) :: restart_time
   real(c_double), allocatable :: dt_rest(:)
   ! steps to the next output, and to the step that reaches totaltime
   integer :: n_batch
   real(dp) :: t_acc

   ! defaults of defaultNamelists (namelists.F90:112-146); dt_max is the
   ! default-real literal 3.e-5 there, kept as such.  tag has no default in
   ! the reference; blank here.
   Nr_max = 20; Nm_max = 20
   eta = 0.35_dp; Ra = 1000.0_dp; Pr = 1.0_dp; mBC = "SF"; ampT = 0.00001_dp
   l_add_pert = .false.; lagpts = 14; buo_tscheme = "IMP"; n_init = 0
   n_time_steps = 100; dt = 0.01_dp; totaltime = 100.0_dp; l_vartimestep = .false.; l_optimizedt = .false.
   time_scheme_type = "IMEXRK"; time_scheme_imp = "ARS222"; time_scheme_exp = "expARS222"
   l_imexrk_started = .true.; dt_coef = 2.0_dp; dt_max = 3.e-5; CFL = 0.9_dp
   l_restart = .false.; n_restart = 0; n_restart_point = 0; n_snapshot_point = 0
   tag = ""; n_checkpoint = 500000; n_snapshot = 300000; n_KE = 1; n_KEspec = 100000
   seed = -1; l_const_history = .false.; solver = "auto"

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
   p%l_restart = merge(1, 0, l_const_history .or. l_restart)
   p%l_vartimestep = merge(1, 0, l_vartimestep)
   p%l_optimizedt = merge(1, 0, l_optimizedt)
   p%CFL = CFL
   p%dt_coef = dt_coef
   p%dt_max = dt_max
   p%lagpts = lagpts
   p%n_KE = n_KE     ! keeps what the power budget and the mean-flow balance read
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

   rc = fdns_scheme_orders(p, n_imp, n_exp)
   if (rc /= FDNS_OK) call fail("fdns_scheme_orders", rc)
   if (l_restart) call read_checkpoints()
   rc = fdns_create(p, run)
   if (rc /= FDNS_OK) call fail("fdns_create", rc)
   if (l_restart) then
      rc = fdns_restart(run, size(tS, 3), c_loc(tS), c_loc(oS), c_loc(rS), c_loc(pS), restart_time)
      if (rc /= FDNS_OK) call fail("fdns_restart", rc)
      rc = fdns_restart_dt(run, size(dt_rest), dt_rest)
      if (rc /= FDNS_OK) call fail("fdns_restart_dt", rc)
      deallocate (tS, oS, rS, pS)
   else if (seed >= 0) then
      rc = fdns_seed_state(run, seed, ampT)
      if (rc /= FDNS_OK) call fail("fdns_seed_state", rc)
   else
      rc = fdns_init_perturbation(run, n_init, ampT)
      if (rc /= FDNS_OK) call fail("fdns_init_perturbation", rc)
   end if

   open (newunit=ke_unit, file="KE_"//trim(tag), status="unknown", position="append", action="write")
   open (newunit=nu_unit, file="Nusselt_"//trim(tag), status="unknown", position="append", action="write")
   open (newunit=pw_unit, file="power_"//trim(tag), status="unknown", position="append", action="write")
   open (newunit=up_unit, file="upbalance_"//trim(tag), status="unknown", position="append", action="write")
   open (newunit=ks_unit, file="KEspec_"//trim(tag), status="unknown", position="append", action="write")
   open (newunit=ts_unit, file="Temp_spec_"//trim(tag), status="unknown", position="append", action="write")
   allocate (upbal(3, Nr_max), kesp(Nm_max + 1), tsp(Nm_max + 1))

   count_chkpnt = n_restart_point
   count_snap = n_snapshot_point
   call system_clock(t0, rate)
   n_step = n_restart
   n_end = n_time_steps + n_restart
   rc = fdns_diag(run, d)
   if (rc /= FDNS_OK) call fail("fdns_diag", rc)
   do while (n_step < n_end)
      ! advance to the next step with output due, in one call
      nxt = min(n_end, next_multiple(n_step, n_KE), next_multiple(n_step, n_checkpoint), &
                next_multiple(n_step, n_snapshot), next_multiple(n_step, n_KEspec))
      ! the IMEX and IMEXRK loops stop after the step whose time reaches
      ! totaltime, writing nothing for it (steptime.F90:255, 287-292;
      ! steptime_imexrk.F90:187-228); the RK loop has no such exit.  With
      ! constant dt that step is found ahead by adding dt as the loop does;
      ! with variable dt the run goes one step at a time.
      if (trim(time_scheme_type) /= "RK") then
         if (l_vartimestep) then
            nxt = n_step + 1
         else
            t_acc = d(1)
            do n_batch = 1, nxt - n_step
               t_acc = t_acc + dt
               if (t_acc >= totaltime) exit
            end do
            nxt = min(nxt, n_step + n_batch)
         end if
      end if
      rc = fdns_step(run, nxt - n_step)
      if (rc /= FDNS_OK) call fail("fdns_step", rc)
      n_step = nxt
      rc = fdns_diag(run, d)
      if (rc /= FDNS_OK) call fail("fdns_diag", rc)
      rc = fdns_get_dt(run, dt_new, c_null_ptr)
      if (rc /= FDNS_OK) call fail("fdns_get_dt", rc)
      if (trim(time_scheme_type) /= "RK" .and. d(1) >= totaltime) exit
      if (due(n_snapshot)) then
         count_snap = count_snap + 1
         call write_snapshot(count_snap, d(1))
      end if
      if (due(n_checkpoint)) then
         count_chkpnt = count_chkpnt + 1
         call write_checkpoint(count_chkpnt, d(1))
      end if
      if (due(n_KE)) then
         rc = fdns_budget(run, bud, c_loc(upbal))
         if (rc /= FDNS_OK) call fail("fdns_budget", rc)
         ! upbalance_<tag>: radius, d u_phi,0/dt, advective and viscous terms (output.F90:224-226)
         do j = 1, Nr_max
            write (up_unit, '(4ES20.12)') radius_at(j), upbal(1, j), upbal(2, j), upbal(3, j)
         end do
         ! KE_<tag>: dt, time, KE, KE of u_r, KE of u_phi, KE of the mean flow (output.F90:256)
         write (ke_unit, '(6es20.12)') dt_new, d(1), d(2), d(5), d(6), bud(1)
         ! Nusselt_<tag>: time, Nu_t, Nu_b (output.F90:265)
         write (nu_unit, '(3ES20.12)') d(1), d(3), d(4)
         ! power_<tag>: time, viscous dissipation, buoyancy power (output.F90:314)
         write (pw_unit, '(3ES20.12)') d(1), bud(2), bud(3)
         if (d(2) > 1.0e10_dp) then
            write (0, '(a)') "kinetic energy above 1e10: stopping (output.F90:321-323)"
            exit
         end if
      end if
      if (due(n_KEspec)) then
         ! KEspec_<tag>, Temp_spec_<tag>: step, m, value (output.F90:350-353)
         rc = fdns_spectra(run, kesp, tsp)
         if (rc /= FDNS_OK) call fail("fdns_spectra", rc)
         do j = 0, Nm_max
            write (ks_unit, '(2i8, 1ES20.12)') n_step, j, kesp(j + 1)
            write (ts_unit, '(2i8, 1ES20.12)') n_step, j, tsp(j + 1)
         end do
      end if
   end do
   call system_clock(t1)

   close (ke_unit)
   close (nu_unit)
   close (pw_unit)
   close (up_unit)
   close (ks_unit)
   close (ts_unit)
   write (6, '(a,i0,a,es12.5,a)') "steps: ", n_step - n_restart, "  wall time [s]: ", &
      & real(t1 - t0, dp) / real(rate, dp), ""
   rc = fdns_destroy(run)

contains

   subroutine fail(what, code)
      character(len=*), intent(in) :: what
      integer(c_int), intent(in) :: code
      write (0, '(a,a,i0,a,a)') what, " failed (", code, "): ", trim(fdns_errmsg(code))
      stop 1
   end subroutine fail

   ! the first multiple of k after n (huge when k <= 0: never)
   integer function next_multiple(n, k)
      integer, intent(in) :: n, k
      if (k <= 0) then
         next_multiple = huge(n)
      else
         next_multiple = (n / k + 1) * k
      end if
   end function next_multiple

   ! radius(i) of init_grid (init.F90:129-131)
   real(dp) function radius_at(i)
      integer, intent(in) :: i
      real(dp) :: pi, rmin
      pi = 4.0_dp * atan(1.0_dp)
      rmin = eta / (1.0_dp - eta)
      radius_at = (rmin + (0.5_dp * (1.0_dp + cos(real(Nr_max - i, kind=dp) * pi / real(Nr_max - 1, kind=dp)))))
   end function radius_at

   logical function due(k)
      integer, intent(in) :: k
      due = k > 0
      if (due) due = mod(n_step, k) == 0
   end function due

   ! store_snapshot (output.F90:488-516, also the RK path's); the IMEXRK
   ! variant adds the omgFR_check record, which that path never assigns
   ! (zeros here)
   subroutine write_snapshot(count, tot_time)
      integer, intent(in) :: count
      real(dp), intent(in) :: tot_time
      complex(dp), allocatable, target :: tF(:,:), oF(:,:), rF(:,:), pF(:,:), zF(:,:)
      real(dp), allocatable :: radius(:)
      character(len=72) :: name
      real(dp) :: pi, rmin
      integer :: i, unit
      allocate (tF(Nm_max+1, Nr_max), oF(Nm_max+1, Nr_max), rF(Nm_max+1, Nr_max), pF(Nm_max+1, Nr_max))
      rc = fdns_get_state(run, c_loc(tF), c_loc(oF), c_loc(rF), c_loc(pF), c_null_ptr)
      if (rc /= FDNS_OK) call fail("fdns_get_state", rc)
      ! init_grid (init.F90:129-131)
      pi = 4.0_dp * atan(1.0_dp)
      rmin = eta / (1.0_dp - eta)
      allocate (radius(Nr_max))
      do i = 1, Nr_max
         radius(i) = (rmin + (0.5_dp * (1.0_dp + cos(real(Nr_max - i, kind=dp) * pi / real(Nr_max - 1, kind=dp)))))
      end do
      write (name, fmt='(a,I5.5)') "snapshot_plot_", count
      open (newunit=unit, file=name, action="write", status="replace", form='unformatted')
      write (unit) Ra, Pr, eta
      write (unit) dt_new, tot_time
      write (unit) Nm_max, Nr_max
      write (unit) radius
      write (unit) tF
      write (unit) oF
      write (unit) rF
      write (unit) pF
      if (trim(time_scheme_type) == "IMEXRK") then
         allocate (zF(Nm_max+1, Nr_max))
         zF = (0.0_dp, 0.0_dp)
         write (unit) zF
      end if
      close (unit)
   end subroutine write_snapshot

   ! store_checkpoint / store_checkpoint_imexrk (output.F90:359-450)
   subroutine write_checkpoint(count, tot_time)
      integer, intent(in) :: count
      real(dp), intent(in) :: tot_time
      complex(dp), allocatable, target :: tF(:,:), oF(:,:), rF(:,:), pF(:,:)
      complex(dp), allocatable, target :: hit(:,:,:), het(:,:,:), hiw(:,:,:), hew(:,:,:), hiu(:,:), heu(:,:)
      real(dp), allocatable, target :: dt_array(:)
      character(len=72) :: name
      integer :: hi, he, n_max, unit
      n_max = max(n_imp, n_exp)
      ! history slots: one on the IMEXRK path (store_checkpoint_imexrk); the
      ! IMEX orders, or the RK path's stage counts, in store_checkpoint
      hi = 1; he = 1
      if (trim(time_scheme_type) /= "IMEXRK") then
         hi = n_imp; he = n_exp
      end if
      allocate (tF(Nm_max+1, Nr_max), oF(Nm_max+1, Nr_max), rF(Nm_max+1, Nr_max), pF(Nm_max+1, Nr_max))
      allocate (hit(hi, Nm_max+1, Nr_max), het(he, Nm_max+1, Nr_max), hiw(hi, Nm_max+1, Nr_max), &
                hew(he, Nm_max+1, Nr_max), hiu(hi, Nr_max), heu(he, Nr_max))
      allocate (dt_array(n_max))
      rc = fdns_get_dt(run, dt_new, c_loc(dt_array))
      if (rc /= FDNS_OK) call fail("fdns_get_dt", rc)
      rc = fdns_get_state(run, c_loc(tF), c_loc(oF), c_loc(rF), c_loc(pF), c_null_ptr)
      if (rc /= FDNS_OK) call fail("fdns_get_state", rc)
      rc = fdns_get_history(run, c_loc(hit), c_loc(het), c_loc(hiw), c_loc(hew), c_loc(hiu), c_loc(heu))
      if (rc /= FDNS_OK) call fail("fdns_get_history", rc)
      write (name, fmt='(a,I5.5)') "checkpoint_", count
      open (newunit=unit, file=name, action="write", status="replace", form='unformatted')
      write (unit) dt_new, tot_time, Nm_max, Nr_max, n_imp, n_exp, n_max
      write (unit) dt_array
      write (unit) tF
      write (unit) oF
      write (unit) rF
      write (unit) pF
      write (unit) hit
      write (unit) het
      write (unit) hiw
      write (unit) hew
      write (unit) hiu
      write (unit) heu
      close (unit)
   end subroutine write_checkpoint

   ! get_checkpoint_data / read_checkpoint (init.F90:447-592, 865-935): the
   ! states of the checkpoints the restart start-up needs, before any device work
   subroutine read_checkpoints()
      real(dp), allocatable :: dt_old(:)
      complex(dp), allocatable, target :: tO(:,:), oO(:,:), rO(:,:), pO(:,:)
      real(dp) :: dt_new
      integer :: K, j, unit, ios, Nm_old, Nr_old, ni_old, ne_old, nmax_old, n_max
      character(len=72) :: name
      K = 1
      if (trim(time_scheme_type) == "IMEX" .and. l_imexrk_started) K = n_imp
      ! dt_array (init.F90:464-612); entries no checkpoint sets keep dt (the
      ! reference leaves them unassigned)
      n_max = max(n_imp, n_exp)
      allocate (dt_rest(n_max))
      dt_rest = dt
      allocate (tS(Nm_max+1, Nr_max, K), oS(Nm_max+1, Nr_max, K), rS(Nm_max+1, Nr_max, K), pS(Nm_max+1, Nr_max, K))
      do j = 1, K
         write (name, '("checkpoint_",I5.5)') n_restart_point - K + j
         open (newunit=unit, status='old', file=name, form='unformatted', action='read', iostat=ios)
         if (ios /= 0) then
            write (0, '(a)') "restart checkpoint not found: "//trim(name)
            stop 2
         end if
         read (unit, iostat=ios) dt_new, restart_time, Nm_old, Nr_old, ni_old, ne_old, nmax_old
         if (ios /= 0) call bad_file(name)
         if (Nm_old /= Nm_max .or. Nr_old /= Nr_max) then
            ! init.F90:722-861: one checkpoint's state on the finer grid
            if (K /= 1 .or. Nm_old > Nm_max .or. Nr_old > Nr_max .or. Nm_old < 1 .or. Nr_old < 4) then
               write (0, '(a,4(a,i0))') trim(name), ": Nm_max=", Nm_old, " Nr_max=", Nr_old, &
                  & " differ from the namelist's ", Nm_max, " x ", Nr_max
               write (0, '(a)') "a change of resolution refines one checkpoint (no multistep histories)"
               stop 2
            end if
         end if
         if (nmax_old < 1 .or. nmax_old > 64) call bad_file(name)
         allocate (dt_old(nmax_old))
         read (unit, iostat=ios) dt_old
         if (ios /= 0) call bad_file(name)
         if (K > 1) then
            ! checkpoint j of K (oldest first) gives entry K - j + 1
            dt_rest(K - j + 1) = dt_old(1)
         else
            ! the newest checkpoint's array, cut or padded with its last entry
            dt_rest(1:min(n_max, nmax_old)) = dt_old(1:min(n_max, nmax_old))
            if (n_max > nmax_old) dt_rest(nmax_old + 1:n_max) = dt_old(nmax_old)
         end if
         deallocate (dt_old)
         allocate (tO(Nm_old+1, Nr_old), oO(Nm_old+1, Nr_old), rO(Nm_old+1, Nr_old), pO(Nm_old+1, Nr_old))
         read (unit, iostat=ios) tO
         if (ios == 0) read (unit, iostat=ios) oO
         if (ios == 0) read (unit, iostat=ios) rO
         if (ios == 0) read (unit, iostat=ios) pO
         if (ios /= 0) call bad_file(name)
         close (unit)
         rc = fdns_refine_state(Nm_old, Nr_old, c_loc(tO), c_loc(oO), c_loc(rO), c_loc(pO), Nm_max, Nr_max, &
                                c_loc(tS(1, 1, j)), c_loc(oS(1, 1, j)), c_loc(rS(1, 1, j)), c_loc(pS(1, 1, j)))
         if (rc /= FDNS_OK) call fail("fdns_refine_state", rc)
         deallocate (tO, oO, rO, pO)
      end do
      if (l_add_pert) tS(:, :, K) = ampT + tS(:, :, K)
   end subroutine read_checkpoints

   subroutine bad_file(name)
      character(len=*), intent(in) :: name
      write (0, '(a)') "restart checkpoint unreadable: "//trim(name)
      stop 2
   end subroutine bad_file

end program fdns_driver
