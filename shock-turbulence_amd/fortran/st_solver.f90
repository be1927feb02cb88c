! st_solver.f90 -- the reference's subroutine interfaces, served by the
! MI355X pipeline through the C-ABI (module st_capi).
!
! The reference selects its operators at link time: every caller names an
! external subroutine (RHS, DFDX, Compact_Filter, LDDRK4, ...) with an
! implicit interface.  The subroutines below keep those names and argument
! lists, so a driver written against the reference links against this file
! instead of src/derivCDopt6_periodic.f, src/filters_periodic.f,
! src/rhsv_periodic_sf.f, src/hyper_rhs.f, src/hyper_dynamic.f and
! src/lddrk4_periodic_sf.f.  The grid extents, which the reference fixes
! at compile time (src/header:5-10), are set once by Setup_Solver.
!
! State: the RK state (/RKWORK/ Q, src/main.f:31) lives on the GPU between
! Upload_State and Download_State; LDDRK4, Compute_CFL and Compute_Tstep
! work on that resident copy (their Q arguments are kept for the
! interface and not read).  The operator entry points (DFDX ... Fix_Pressure,
! RHS, Hyper_RHS) take and return host arrays.  Errors stop the program
! with the library's message, as the reference's fatal paths do
! (src/pentadiagonal.f90:55).
module st_state
  use st_capi
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: Mx = 0, My = 0, Mz = 0
  real(c_double), save :: Tstep = 0.0d0       ! /TIME_RK/ Tstep (src/lddrk4_periodic_sf.f:74-145)
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
  end subroutine Setup_Solver

  subroutine Finalize_Solver()
    if (c_associated(ctx)) call st_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine Finalize_Solver

  subroutine Upload_State(Q)
    real(c_double), intent(in) :: Q(*)
    call st_check(st_state_upload(ctx, Q), 'Upload_State')
  end subroutine Upload_State

  subroutine Download_State(Q)
    real(c_double), intent(inout) :: Q(*)
    call st_check(st_state_download(ctx, Q), 'Download_State')
  end subroutine Download_State
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

! Compute_CFL(Q, cfl_stab, cfl_resv, step)  src/lddrk4_periodic_sf.f:7
subroutine Compute_CFL(Q, cfl_stab, cfl_resv, step)
  use st_state
  real(8), intent(in) :: Q(*)
  real(8), intent(out) :: cfl_stab, cfl_resv
  integer, intent(in) :: step
  call st_check(st_cfl(ctx, cfl_stab, cfl_resv), 'Compute_CFL')
end subroutine Compute_CFL

! Compute_Tstep(Q, cfl_resv, step)  src/lddrk4_periodic_sf.f:74; sets Tstep
subroutine Compute_Tstep(Q, cfl_resv, step)
  use st_state
  real(8), intent(in) :: Q(*)
  real(8), intent(in) :: cfl_resv
  integer, intent(in) :: step
  call st_check(st_tstep(ctx, cfl_resv, int(step, c_int), Tstep), 'Compute_Tstep')
end subroutine Compute_Tstep

! LDDRK4(time, step)  src/lddrk4_periodic_sf.f:311, with the Tstep of the
! last Compute_Tstep
subroutine LDDRK4(time, step)
  use st_state
  real(8), intent(inout) :: time
  integer, intent(inout) :: step
  integer(c_int) :: s
  s = int(step, c_int)
  call st_check(st_step(ctx, Tstep, time, s), 'LDDRK4')
  step = int(s)
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
