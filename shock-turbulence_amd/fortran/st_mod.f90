! st_mod.f90 -- iso_c_binding interfaces to the C-ABI of include/st_api.h.
!
! The derived types mirror st_grid and st_params field for field (C layout);
! every function returns 0 on success and a negative code on failure, with
! the message from st_last_error.  Arrays are Fortran-ordered Q(My,Mz,Mx,5)
! as in the reference (src/main.f:28), passed by address.
module st_capi
  use iso_c_binding
  implicit none

  type, bind(C) :: st_grid
     integer(c_int) :: nx, ny, nz
     real(c_double) :: period_x, period_y, period_z
     real(c_double) :: x_start, y_start, z_start
     integer(c_int) :: nonperiodic_x = 0
  end type st_grid

  ! The input deck of Define_Parameters (src/define_periodic_sf.f:20-80).
  type, bind(C) :: st_deck
     character(kind=c_char) :: vis_path(24), restart_path(24), stats_path(24), probe_path(24), &
          source_path(24), prefix(24)
     integer(c_int) :: grid_x, grid_y
     real(c_double) :: x_start, y_start, z_start, period_x, period_y, period_z
     real(c_double) :: Re, Pr, Sc, gamma, vis_exp, bulk_ratio, T_ratio
     integer(c_int) :: hyper_param, part_param, march_mode
     real(c_double) :: Tstep, cfl_cal
     integer(c_int) :: Nsteps
     real(c_double) :: run_time
     integer(c_int) :: comp_stats
     integer(c_int) :: out_grid, out_init, out_incr_x, out_incr_y, out_incr_z, out_restart, out_vis, &
          out_probe, out_stats, out_source
  end type st_deck

  type, bind(C) :: st_params
     real(c_double) :: Re, Pr, gamma, vis_exp, bulk_ratio
     integer(c_int) :: hyper_param
     real(c_double) :: cfl_cal
     integer(c_int) :: d2_mode
     integer(c_int) :: rk_scheme
  end type st_params

  interface
     integer(c_int) function st_ctx_create(grid, params, ctx) bind(C, name='st_ctx_create')
       import :: c_int, c_ptr, st_grid, st_params
       type(st_grid), intent(in) :: grid
       type(st_params), intent(in) :: params
       type(c_ptr), intent(out) :: ctx
     end function st_ctx_create

     subroutine st_ctx_destroy(ctx) bind(C, name='st_ctx_destroy')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine st_ctx_destroy

     type(c_ptr) function st_last_error(ctx) bind(C, name='st_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function st_last_error

     integer(c_int) function st_sync(ctx) bind(C, name='st_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function st_sync

     integer(c_int) function st_state_upload(ctx, Q) bind(C, name='st_state_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: Q(*)
     end function st_state_upload

     integer(c_int) function st_state_download(ctx, Q) bind(C, name='st_state_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: Q(*)
     end function st_state_download

     integer(c_int) function st_d1(ctx, dir, F, dF, iadd, cadd) bind(C, name='st_d1')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: dir, iadd
       real(c_double), intent(in) :: F(*)
       real(c_double), intent(inout) :: dF(*)
       real(c_double), value :: cadd
     end function st_d1

     integer(c_int) function st_d2(ctx, dir, F, dF, d2F, iadd) bind(C, name='st_d2')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: dir, iadd
       real(c_double), intent(in) :: F(*)
       real(c_double), intent(inout) :: dF(*), d2F(*)
     end function st_d2

     integer(c_int) function st_filter(ctx, F, F0, nf) bind(C, name='st_filter')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: F(*)
       real(c_double), intent(inout) :: F0(*)
       integer(c_int), value :: nf
     end function st_filter

     integer(c_int) function st_gauss(ctx, F, Fg, nf) bind(C, name='st_gauss')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: F(*)
       real(c_double), intent(inout) :: Fg(*)
       integer(c_int), value :: nf
     end function st_gauss

     integer(c_int) function st_rhs(ctx, Q, Qrhs) bind(C, name='st_rhs')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: Q(*)
       real(c_double), intent(inout) :: Qrhs(*)
     end function st_rhs

     integer(c_int) function st_hyper_rhs(ctx, Qrhs) bind(C, name='st_hyper_rhs')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: Qrhs(*)
     end function st_hyper_rhs

     integer(c_int) function st_compute_hyper(ctx, Q) bind(C, name='st_compute_hyper')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: Q(*)
     end function st_compute_hyper

     integer(c_int) function st_get_hyper(ctx, H3) bind(C, name='st_get_hyper')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: H3(*)
     end function st_get_hyper

     integer(c_int) function st_compute_grad(ctx, maxdil, enstrophy) bind(C, name='st_compute_grad')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: maxdil, enstrophy
     end function st_compute_grad

     integer(c_int) function st_fix_pressure(ctx, Q) bind(C, name='st_fix_pressure')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: Q(*)
     end function st_fix_pressure

     integer(c_int) function st_cfl(ctx, cfl_stab, cfl_resv) bind(C, name='st_cfl')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: cfl_stab, cfl_resv
     end function st_cfl

     integer(c_int) function st_tstep(ctx, cfl_resv, step, dt) bind(C, name='st_tstep')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: cfl_resv
       integer(c_int), value :: step
       real(c_double), intent(out) :: dt
     end function st_tstep

     integer(c_int) function st_step(ctx, dt, time, step) bind(C, name='st_step')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       real(c_double), intent(inout) :: time
       integer(c_int), intent(inout) :: step
     end function st_step

     integer(c_int) function st_diagnostics(ctx, out7) bind(C, name='st_diagnostics')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out7(7)
     end function st_diagnostics

     integer(c_int) function st_ic_isotropic(nx, ny, nz, period, k0, urms, mt, gamma, seed, Q) &
          bind(C, name='st_ic_isotropic')
       import :: c_int, c_double, c_long_long
       integer(c_int), value :: nx, ny, nz
       real(c_double), intent(in) :: period(3)
       real(c_double), value :: k0, urms, mt, gamma
       integer(c_long_long), value :: seed
       real(c_double), intent(inout) :: Q(*)
     end function st_ic_isotropic

     integer(c_int) function st_restart_write(path, nx, ny, nz, Q, time, step) bind(C, name='st_restart_write')
       import :: c_int, c_double, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: nx, ny, nz
       real(c_double), intent(in) :: Q(*)
       real(c_double), value :: time
       integer(c_int), value :: step
     end function st_restart_write

     integer(c_int) function st_restart_read(path, nx, ny, nz, Q, time, step) bind(C, name='st_restart_read')
       import :: c_int, c_double, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: nx, ny, nz
       real(c_double), intent(inout) :: Q(*)
       real(c_double), intent(out) :: time
       integer(c_int), intent(out) :: step
     end function st_restart_read

     integer(c_int) function st_deck_read(path, deck) bind(C, name='st_deck_read')
       import :: c_int, c_char, st_deck
       character(kind=c_char), intent(in) :: path(*)
       type(st_deck), intent(out) :: deck
     end function st_deck_read
  end interface

contains

  ! A Fortran string as a NUL-terminated C string.
  function c_str(f) result(c)
    character(len=*), intent(in) :: f
    character(kind=c_char, len=len_trim(f) + 1) :: c
    c = trim(f) // c_null_char
  end function c_str

  ! A NUL-terminated C character array as a Fortran string.
  function f_str(c) result(f)
    character(kind=c_char), intent(in) :: c(:)
    character(len=size(c)) :: f
    integer :: i
    f = ' '
    do i = 1, size(c)
       if (c(i) == c_null_char) exit
       f(i:i) = c(i)
    end do
  end function f_str

  ! Message of the last failure on ctx, as a Fortran string.
  function st_error_text(ctx) result(msg)
    type(c_ptr), value :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = st_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function st_error_text

end module st_capi
