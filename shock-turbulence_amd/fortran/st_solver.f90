! st_solver.f90 -- the reference's subroutine interfaces, served by the
! MI355X pipeline through the C-ABI (module st_capi).
!
! The reference selects its operators at link time: every caller names an
! external subroutine (RHS, DFDX, Compact_Filter, LDDRK4, ...) with an
! implicit interface.  The subroutines below keep those names and argument
! lists, so a driver written against the reference links against this file
! instead of src/derivCDopt6_periodic.f, src/filters_periodic.f,
! src/rhsv_periodic_sf.f, src/hyper_rhs.f, src/hyper_dynamic.f, src/grad.f
! and src/lddrk4_periodic_sf.f.  The grid extents, which the reference
! fixes at compile time (src/header:5-10), are set once by Setup_Solver.
!
! State: the reference keeps the RK state in COMMON /RKWORK/ Q
! (src/main.f:31).  Here Upload_State(Q) copies Q to the GPU and registers
! that host array as /RKWORK/.  LDDRK4 advances the GPU copy and then
! copies it back into the registered array, so the host Q is current after
! every step, as the reference's is (Output and restart writers read it).
! Compute_CFL, Compute_Tstep and Compute_hyper read their Q argument: it
! is copied to the GPU as the state before the reduction, so edits a
! caller makes to Q between steps are honoured.  Their Q must be the
! registered array (a different array stops the program: the reference
! always passes /RKWORK/ Q there).  These copies cost one host transfer of
! the state each; a caller that keeps the state on the GPU uses the C-ABI
! (st_cfl, st_tstep, st_step, st_march) instead.  The operator entry
! points (DFDX ... Fix_Pressure, RHS, Hyper_RHS) take and return host
! arrays.  Errors stop the program with the library's message, as the
! reference's fatal paths do (src/pentadiagonal.f90:55).
module st_state
  use st_capi
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: Mx = 0, My = 0, Mz = 0
  integer, save :: hyper_on = 0, nonper = 0
  real(c_double), save :: Tstep = 0.0d0       ! /TIME_RK/ Tstep (src/lddrk4_periodic_sf.f:74-145)
  real(c_double), save :: MaxDil = 0.0d0, Enstrophy = 0.0d0   ! /TRACK/ (src/header:225-228)
  type(c_ptr), save :: rkwork = c_null_ptr    ! the host array registered as /RKWORK/ Q
contains
  subroutine st_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
       write(*, '(a,a,i4,a,a)') where, ': error ', rc, ': ', trim(st_error_text(ctx))
       error stop 1
    end if
  end subroutine st_check

  ! Setup_Grid + Setup_Derivatives + Init_Filters + Initialize_MPI
  ! (src/grid.f:15, src/derivCDopt6_periodic.f:9, src/filters_periodic.f:9,
  ! src/initialize_mpi_periodic.f90:1) for one GPU.
  subroutine Setup_Solver(grid, params)
    type(st_grid), intent(in) :: grid
    type(st_params), intent(in) :: params
    if (c_associated(ctx)) call st_ctx_destroy(ctx)
    call st_check(st_ctx_create(grid, params, ctx), 'Setup_Solver')
    Mx = grid%nx
    My = grid%ny
    Mz = grid%nz
    hyper_on = params%hyper_param
    nonper = grid%nonperiodic_x
    rkwork = c_null_ptr
  end subroutine Setup_Solver

  subroutine Finalize_Solver()
    if (c_associated(ctx)) call st_ctx_destroy(ctx)
    ctx = c_null_ptr
    rkwork = c_null_ptr
  end subroutine Finalize_Solver

  ! Copies Q to the GPU and registers it as /RKWORK/ Q.
  subroutine Upload_State(Q)
    real(c_double), intent(in), target :: Q(*)
    call st_check(st_state_upload(ctx, Q), 'Upload_State')
    rkwork = c_loc(Q(1))
  end subroutine Upload_State

  subroutine Download_State(Q)
    real(c_double), intent(inout) :: Q(*)
    call st_check(st_state_download(ctx, Q), 'Download_State')
  end subroutine Download_State

  ! Q as the state: it must be the registered /RKWORK/ array (or the first
  ! array seen, which is then registered).
  subroutine check_state(Q, where)
    real(c_double), intent(in), target :: Q(*)
    character(len=*), intent(in) :: where
    if (.not. c_associated(rkwork)) rkwork = c_loc(Q(1))
    if (.not. c_associated(rkwork, c_loc(Q(1)))) then
       write(*, '(a,a)') where, ': Q is not the state array registered by Upload_State (/RKWORK/ Q)'
       error stop 1
    end if
  end subroutine check_state

  ! check_state, then Q is copied to the GPU as the state.
  subroutine read_state(Q, where)
    real(c_double), intent(in), target :: Q(*)
    character(len=*), intent(in) :: where
    call check_state(Q, where)
    call st_check(st_state_upload(ctx, Q), where)
  end subroutine read_state

  ! Copies the GPU state back into the registered /RKWORK/ array.
  subroutine write_back()
    real(c_double), pointer :: q(:)
    if (.not. c_associated(rkwork)) return
    call c_f_pointer(rkwork, q, [5 * Mx * My * Mz])
    call st_check(st_state_download(ctx, q), 'LDDRK4')
  end subroutine write_back
end module st_state

! DFDX/DFDY/DFDZ(F, dF, iadd, cadd)  src/derivCDopt6_periodic.f:164,400,631
subroutine DFDX(F, dF, iadd, cadd)
  use st_state
  real(8), intent(in) :: F(*)
  real(8), intent(inout) :: dF(*)
  integer, intent(in) :: iadd
  real(8), intent(in) :: cadd
  call st_check(st_d1(ctx, 0_c_int, F, dF, int(iadd, c_int), cadd), 'DFDX')
end subroutine DFDX

subroutine DFDY(F, dF, iadd, cadd)
  use st_state
  real(8), intent(in) :: F(*)
  real(8), intent(inout) :: dF(*)
  integer, intent(in) :: iadd
  real(8), intent(in) :: cadd
  call st_check(st_d1(ctx, 1_c_int, F, dF, int(iadd, c_int), cadd), 'DFDY')
end subroutine DFDY

subroutine DFDZ(F, dF, iadd, cadd)
  use st_state
  real(8), intent(in) :: F(*)
  real(8), intent(inout) :: dF(*)
  integer, intent(in) :: iadd
  real(8), intent(in) :: cadd
  call st_check(st_d1(ctx, 2_c_int, F, dF, int(iadd, c_int), cadd), 'DFDZ')
end subroutine DFDZ

! D2FDX2/D2FDY2/D2FDZ2(F, dF, d2F, iadd)  src/derivCDopt6_periodic.f:281,515,747
subroutine D2FDX2(F, dF, d2F, iadd)
  use st_state
  real(8), intent(in) :: F(*)
  real(8), intent(inout) :: dF(*), d2F(*)
  integer, intent(in) :: iadd
  call st_check(st_d2(ctx, 0_c_int, F, dF, d2F, int(iadd, c_int)), 'D2FDX2')
end subroutine D2FDX2

subroutine D2FDY2(F, dF, d2F, iadd)
  use st_state
  real(8), intent(in) :: F(*)
  real(8), intent(inout) :: dF(*), d2F(*)
  integer, intent(in) :: iadd
  call st_check(st_d2(ctx, 1_c_int, F, dF, d2F, int(iadd, c_int)), 'D2FDY2')
end subroutine D2FDY2

subroutine D2FDZ2(F, dF, d2F, iadd)
  use st_state
  real(8), intent(in) :: F(*)
  real(8), intent(inout) :: dF(*), d2F(*)
  integer, intent(in) :: iadd
  call st_check(st_d2(ctx, 2_c_int, F, dF, d2F, int(iadd, c_int)), 'D2FDZ2')
end subroutine D2FDZ2

! Compact_Filter(F, F0, N)  src/filters_periodic.f:90 (F0 may be F)
subroutine Compact_Filter(F, F0, N)
  use st_state
  real(8), intent(in) :: F(*)
  real(8), intent(inout) :: F0(*)
  integer, intent(in) :: N
  call st_check(st_filter(ctx, F, F0, int(N, c_int)), 'Compact_Filter')
end subroutine Compact_Filter

! Gaussian_Filter(F, Fg, N)  src/filters_periodic.f:540
subroutine Gaussian_Filter(F, Fg, N)
  use st_state
  real(8), intent(in) :: F(*)
  real(8), intent(inout) :: Fg(*)
  integer, intent(in) :: N
  call st_check(st_gauss(ctx, F, Fg, int(N, c_int)), 'Gaussian_Filter')
end subroutine Gaussian_Filter

! RHS(Q, Qrhs)  src/rhsv_periodic_sf.f:9 (Q is not modified)
subroutine RHS(Q, Qrhs)
  use st_state
  real(8), intent(in) :: Q(*)
  real(8), intent(inout) :: Qrhs(*)
  call st_check(st_rhs(ctx, Q, Qrhs), 'RHS')
end subroutine RHS

! Hyper_RHS(Q, Qrhs)  src/hyper_rhs.f:12; Q must be the argument of the
! preceding RHS call (the gradients it left in /GRAD/ are reused).
subroutine Hyper_RHS(Q, Qrhs)
  use st_state
  real(8), intent(in) :: Q(*)
  real(8), intent(inout) :: Qrhs(*)
  call st_check(st_hyper_rhs(ctx, Qrhs), 'Hyper_RHS')
end subroutine Hyper_RHS

! Fix_Pressure(Q)  src/lddrk4_periodic_sf.f:436
subroutine Fix_Pressure(Q)
  use st_state
  real(8), intent(inout) :: Q(*)
  call st_check(st_fix_pressure(ctx, Q), 'Fix_Pressure')
end subroutine Fix_Pressure

! Compute_hyper(Q)  src/hyper_dynamic.f:173 (src/hyper_routines.f:66 with a
! non-periodic x), as Program Main calls it before the first step
! (src/main.f:100): reads Q and the /GRAD/ of the last RHS or Compute_Grad.
subroutine Compute_hyper(Q)
  use st_state
  real(8), intent(in), target :: Q(*)
  call check_state(Q, 'Compute_hyper')
  call st_check(st_compute_hyper(ctx, Q), 'Compute_hyper')
end subroutine Compute_hyper

! The artificial-transport coefficients of /HYPER/ (beta_hyper, visc_hyper,
! k_hyper, src/hyper_dynamic.f:201-202) as the last Hyper_RHS or
! Compute_hyper left them, for callers that read that COMMON block.
subroutine Get_Hyper(beta_hyper, visc_hyper, k_hyper)
  use st_state
  real(8), intent(out) :: beta_hyper(*), visc_hyper(*), k_hyper(*)
  real(c_double), allocatable :: h(:)
  integer :: n
  n = Mx * My * Mz
  allocate(h(3 * n))
  call st_check(st_get_hyper(ctx, h), 'Get_Hyper')
  beta_hyper(1:n) = h(1:n)
  visc_hyper(1:n) = h(n + 1:2 * n)
  k_hyper(1:n) = h(2 * n + 1:3 * n)
end subroutine Get_Hyper

! Compute_Grad  src/grad.f:13-115 (from Output, src/out.f:718) on the state:
! the velocity gradients of /GRAD/, MaxDil and Enstrophy (/TRACK/)
subroutine Compute_Grad()
  use st_state
  call st_check(st_compute_grad(ctx, MaxDil, Enstrophy), 'Compute_Grad')
end subroutine Compute_Grad

! Compute_CFL(Q, cfl_stab, cfl_resv, step)  src/lddrk4_periodic_sf.f:7; reads Q
subroutine Compute_CFL(Q, cfl_stab, cfl_resv, step)
  use st_state
  real(8), intent(in), target :: Q(*)
  real(8), intent(out) :: cfl_stab, cfl_resv
  integer, intent(in) :: step
  call read_state(Q, 'Compute_CFL')
  call st_check(st_cfl(ctx, cfl_stab, cfl_resv), 'Compute_CFL')
end subroutine Compute_CFL

! Compute_Tstep(Q, cfl_resv, step)  src/lddrk4_periodic_sf.f:74; sets Tstep.
! Q is read where the reference reads it: for the artificial-transport
! limits (hyper_param = 1 and step > 1, :99; from the first step with a
! non-periodic x, src/lddrk4_nonperiodic_sf.f:105).
subroutine Compute_Tstep(Q, cfl_resv, step)
  use st_state
  real(8), intent(in), target :: Q(*)
  real(8), intent(in) :: cfl_resv
  integer, intent(in) :: step
  if (hyper_on == 1 .and. (step > 1 .or. nonper == 1)) call read_state(Q, 'Compute_Tstep')
  call st_check(st_tstep(ctx, cfl_resv, int(step, c_int), Tstep), 'Compute_Tstep')
end subroutine Compute_Tstep

! LDDRK4(time, step)  src/lddrk4_periodic_sf.f:311, with the Tstep of the
! last Compute_Tstep; the registered /RKWORK/ Q receives the new state
subroutine LDDRK4(time, step)
  use st_state
  real(8), intent(inout) :: time
  integer, intent(inout) :: step
  integer(c_int) :: s
  s = int(step, c_int)
  call st_check(st_step(ctx, Tstep, time, s), 'LDDRK4')
  step = int(s)
  call write_back()
end subroutine LDDRK4

! Compute_Mom, Compute_KE, Compute_Mass, Compute_Energy, Compute_Entropy
! (src/lddrk4_periodic_sf.f:153-302) in one call
subroutine Compute_Diagnostics(MomX, MomY, MomZ, KE, Mass_t, e_t, Entropy)
  use st_state
  real(8), intent(out) :: MomX, MomY, MomZ, KE, Mass_t, e_t, Entropy
  real(c_double) :: d(7)
  call st_check(st_diagnostics(ctx, d), 'Compute_Diagnostics')
  MomX = d(1)
  MomY = d(2)
  MomZ = d(3)
  KE = d(4)
  Mass_t = d(5)
  e_t = d(6)
  Entropy = d(7)
end subroutine Compute_Diagnostics
