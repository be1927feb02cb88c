! st_run.f90 -- driver of the periodic solver on one MI355X, following
! Program Main (src/main.f:59-184): set up, initial condition, Compute_hyper,
! Compute_CFL / Compute_Tstep, then per step (march mode 1, constant CFL)
! Compute_CFL, Compute_Tstep, LDDRK4, Compute_Grad (the part of Output the
! log needs) and one line of the 'output' log; march mode 2 (constant time
! step) keeps the start-up Tstep.  Modes 3 and 4 run to a wall-clock
! budget and are refused.
!
! Input: a namelist file named by the first argument (defaults below).
! The initial condition is the seeded isotropic field of st_ic_isotropic
! (the reference restarts from a file instead, src/main.f:69-71).  With
! Re <= 0 the Reynolds number is set from re_lambda at t = 0:
! Re = (1/M_t^2)^vis_exp / (u_rms (2/k0) / Re_lambda).
!
! With deck set, the run settings come from an input deck in the
! reference's format (src/define_periodic_sf.f:20-80): periods, starts, Re,
! Pr, gamma, vis_exp, bulk_ratio, hyper_param, cfl_cal and Nsteps; and, as
! in src/main.f:69-71, the state starts from <restart_path>/<prefix>.restart.0
! when that file exists.  restart_in names a restart file to start from
! explicitly.  With restart_every > 0, restart files
! <restart_dir>/<prefix>.restart.<n>, n = 1, 2, ... are written every that
! many steps (src/out.f:722-725; the reference writes every 50), in the
! format of Write_Restart (src/out.f:493-591).
!
! At start-up, as src/main.f:96-104: with out_init = 1, Compute_Grad fills
! the velocity gradients of /GRAD/ first (Output(time, 0)); Compute_hyper
! then sets the artificial-transport coefficients that the first dt limits
! read (they bind from the first step after a restart at step > 1, and from
! step 1 with a non-periodic x).
!
! Output: the 12 value columns of src/main.f:109-116 (n, t, dt, CFL_resv,
! KE, MaxDil, Enstrophy, MomX, MomY, MomZ, Mass, Energy, Entropy) in the
! reference's formats 10 and 11 (:316-317), appended to the file 'output',
! and the same values at full precision on standard output.  With qout
! set, the final state is written there as raw FP64 Q(My,Mz,Mx,5).
program st_run
  use st_capi
  use st_state
  implicit none
  integer :: nx = 64, ny = 64, nz = 64, nsteps = 10, hyper_param = 1, d2_mode = 0, rk_scheme = 0
  integer(8) :: seed = 12345
  real(8) :: k0 = 4.0d0, urms = 1.0d0, mt = 0.3d0, re_lambda = 40.0d0, Re = 0.0d0, Pr = 0.7d0
  real(8) :: gamma = 1.4d0, vis_exp = 0.76d0, bulk_ratio = 0.0d0, cfl_cal = 0.5d0
  real(8) :: period = 6.283185307179586d0
  character(len=256) :: qout = ' ', deck = ' ', restart_in = ' ', restart_dir = '.', prefix = 'st'
  integer :: restart_every = 0, march_mode = 1, out_init = 0
  namelist /run/ nx, ny, nz, nsteps, hyper_param, d2_mode, rk_scheme, seed, k0, urms, mt, &
       re_lambda, Re, Pr, gamma, vis_exp, bulk_ratio, cfl_cal, period, qout, deck, restart_in, &
       restart_dir, prefix, restart_every, march_mode, out_init

  type(st_grid) :: grid
  type(st_params) :: params
  real(8), allocatable, target :: Q(:, :, :, :)
  real(8) :: time, cfl_stab, cfl_resv, per(3), start(3)
  real(8) :: MomX, MomY, MomZ, KE, Mass_t, e_t, Entropy
  integer :: step, step_count, u, ios, cue
  character(len=256) :: arg, fname
  type(st_deck) :: dk
  logical :: have

  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg)
     open(newunit=u, file=trim(arg), status='old', action='read', iostat=ios)
     if (ios /= 0) then
        write(*, '(a,a)') 'st_run: cannot open ', trim(arg)
        error stop 2
     end if
     read(u, nml=run)
     close(u)
  end if
  per = period
  start = 0.0d0
  if (len_trim(deck) > 0) then
     call st_check(st_deck_read(c_str(deck), dk), 'st_deck_read ' // trim(deck))
     per = [dk%period_x, dk%period_y, dk%period_z]
     start = [dk%x_start, dk%y_start, dk%z_start]
     Re = dk%Re
     Pr = dk%Pr
     gamma = dk%gamma
     vis_exp = dk%vis_exp
     bulk_ratio = dk%bulk_ratio
     hyper_param = dk%hyper_param
     cfl_cal = dk%cfl_cal
     nsteps = dk%Nsteps
     march_mode = dk%march_mode
     out_init = dk%out_init
     prefix = f_str(dk%prefix)
     restart_dir = f_str(dk%restart_path)
     if (len_trim(restart_in) == 0) then
        fname = trim(restart_dir) // '/' // trim(prefix) // '.restart.0'
        inquire(file=trim(fname), exist=have)
        if (have) restart_in = fname
     end if
  else if (Re <= 0.0d0) then
     Re = (1.0d0 / (mt * mt)) ** vis_exp / (urms * (2.0d0 / k0) / re_lambda)
  end if

  if (march_mode /= 1 .and. march_mode /= 2) then
     write(*, '(a,i0,a)') 'st_run: march_mode ', march_mode, &
          ' (a wall-clock run_time budget, src/main.f:186-260) is not supported; use 1 or 2'
     error stop 2
  end if

  grid = st_grid(nx, ny, nz, per(1), per(2), per(3), start(1), start(2), start(3))
  params = st_params(Re, Pr, gamma, vis_exp, bulk_ratio, hyper_param, cfl_cal, d2_mode, rk_scheme)
  call Setup_Solver(grid, params)

  allocate(Q(ny, nz, nx, 5))
  time = 0.0d0
  step = 0
  if (len_trim(restart_in) > 0) then
     call st_check(st_restart_read(c_str(restart_in), nx, ny, nz, Q, time, step), &
          'st_restart_read ' // trim(restart_in))
  else
     call st_check(st_ic_isotropic(nx, ny, nz, per, k0, urms, mt, gamma, seed, Q), 'st_ic_isotropic')
  end if
  call Upload_State(Q)
  cue = 1
  call Compute_Diagnostics(MomX, MomY, MomZ, KE, Mass_t, e_t, Entropy)
  if (out_init == 1) call Compute_Grad()                 ! Output(time, 0), src/main.f:96
  call Compute_hyper(Q)                                  ! src/main.f:100
  call Compute_CFL(Q, cfl_stab, cfl_resv, step)
  call Compute_Tstep(Q, cfl_resv, step)
  call log_header()
  call log_line()

  do step_count = 1, nsteps
     if (march_mode == 1) then                           ! src/main.f:129-136
        call Compute_CFL(Q, cfl_stab, cfl_resv, step)
        call Compute_Tstep(Q, cfl_resv, step)
     else                                                ! src/main.f:158-168: the start-up Tstep
        call Compute_CFL(Q, cfl_stab, cfl_resv, step)
        cfl_stab = cfl_stab * Tstep
        cfl_resv = cfl_resv * Tstep
     end if
     call LDDRK4(time, step)
     call Compute_Grad()                                 ! Output(time, step), src/out.f:718
     call Compute_Diagnostics(MomX, MomY, MomZ, KE, Mass_t, e_t, Entropy)
     call log_line()
     if (restart_every > 0) then
        if (mod(step, restart_every) == 0) then
           write(fname, '(a,a,a,a,i0)') trim(restart_dir), '/', trim(prefix), '.restart.', cue
           call st_check(st_restart_write(c_str(fname), nx, ny, nz, Q, time, step), &
                'st_restart_write ' // trim(fname))
           cue = cue + 1
        end if
     end if
  end do

  if (len_trim(qout) > 0) then
     open(newunit=u, file=trim(qout), access='stream', form='unformatted', status='replace')
     write(u) Q
     close(u)
  end if
  call Finalize_Solver()

contains

  ! formats 10 and 11 of src/main.f:316-317; the header of :109-112
  subroutine log_header()
    open(newunit=u, file='output', form='formatted', status='unknown', position='append')
    write(u, '(1x,a5,12(a13))') 'n ', 't   ', 'dt    ', 'CFL_resv  ', 'KE     ', 'MaxDil  ', &
         'Enstrophy ', 'MomX  ', 'MomY  ', 'MomZ  ', 'Mass  ', 'Energy ', 'Entropy'
    close(u)
  end subroutine log_header

  subroutine log_line()
    open(newunit=u, file='output', form='formatted', status='unknown', position='append')
    write(u, '(1x,i5,12(ES13.5))') step, time, Tstep, cfl_resv * Tstep, KE, MaxDil, Enstrophy, &
         MomX, MomY, MomZ, Mass_t, e_t, Entropy
    close(u)
    write(*, '(a,i6,12(1x,ES24.16E3))') 'step', step, time, Tstep, cfl_resv * Tstep, KE, MaxDil, &
         Enstrophy, MomX, MomY, MomZ, Mass_t, e_t, Entropy
  end subroutine log_line

end program st_run
