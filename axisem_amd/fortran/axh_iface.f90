!=========================================================================================
!> iso_c_binding interfaces of the GPU time-loop library (include/axihip.h) and of the
!! host glue (axh_host_glue.c).  Used by the replacement time_evol_wave module and by the
!! GPU convolution of the post-processing program (convolve_gpu.f90).
module axh_iface

  use, intrinsic :: iso_c_binding
  implicit none
  public

  integer(c_int32_t), parameter :: AXH_ABI_VERSION = 1
  integer(c_int), parameter :: AXH_OK = 0, AXH_ERR_BLOWUP = -6

  integer(c_int), parameter :: AXH_MONOPOLE = 0, AXH_DIPOLE = 1, AXH_QUADPOLE = 2
  integer(c_int), parameter :: AXH_NEWMARK2 = 0, AXH_SYMPLEC4 = 1, AXH_ML_SO4M5 = 2, &
                               AXH_ML_SO6M7 = 3

  ! enum axh_solid_term
  integer(c_int), parameter :: AXH_M11S = 0, AXH_M21S = 1, AXH_M41S = 2, AXH_M12S = 3, &
       AXH_M22S = 4, AXH_M42S = 5, AXH_M11Z = 6, AXH_M21Z = 7, AXH_M41Z = 8, &
       AXH_M32S = 9, AXH_M13S = 10, AXH_M33S = 11, AXH_M43S = 12, AXH_M1PHI = 13, &
       AXH_M2PHI = 14, AXH_M4PHI = 15, AXH_M_1 = 16, AXH_M_2 = 17, AXH_M_3 = 18, &
       AXH_M_4 = 19, AXH_M_5 = 20, AXH_M_6 = 21, AXH_M_7 = 22, AXH_M_8 = 23, &
       AXH_M_W1 = 24, AXH_M_W2 = 25, AXH_M_W3 = 26, AXH_M_W4 = 27, AXH_M_W5 = 28, &
       AXH_M0_W1 = 64
  ! enum axh_fluid_term
  integer(c_int), parameter :: AXH_M1CHI_FL = 0, AXH_M2CHI_FL = 1, AXH_M4CHI_FL = 2, &
       AXH_M_W_FL = 3, AXH_M0_W_FL = 64

  type, bind(C) :: axh_dims_t
     integer(c_int32_t) :: abi_version, npol, nel_solid, nel_fluid
     integer(c_int32_t) :: nglob_solid, nglob_fluid, nel_bdry
     integer(c_int32_t) :: src_type, scheme, anel, coarse_grained, n_sls
     integer(c_int32_t) :: niter, seis_it, num_rec, nelsrc, fluid_src
     integer(c_int32_t) :: reserved(5)
     real(c_double)     :: deltat, magnitude
  end type

  interface
     integer(c_int) function axh_create(dims, ctx) bind(C, name='axh_create')
       import :: c_int, c_ptr, axh_dims_t
       type(axh_dims_t), intent(in) :: dims
       type(c_ptr), intent(out)     :: ctx
     end function

     integer(c_int) function axh_destroy(ctx) bind(C, name='axh_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function

     type(c_ptr) function axh_last_error() bind(C, name='axh_last_error')
       import :: c_ptr
     end function

     integer(c_int) function axh_upload_spectral(ctx, G0, G1, G1T, G2, G2T) &
          bind(C, name='axh_upload_spectral')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: G0(*), G1(*), G1T(*), G2(*), G2T(*)
     end function

     integer(c_int) function axh_upload_solid_term(ctx, term, dat) &
          bind(C, name='axh_upload_solid_term')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       integer(c_int), value :: term
       real(c_float), intent(in) :: dat(*)
     end function

     integer(c_int) function axh_upload_fluid_term(ctx, term, dat) &
          bind(C, name='axh_upload_fluid_term')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       integer(c_int), value :: term
       real(c_float), intent(in) :: dat(*)
     end function

     integer(c_int) function axh_upload_mass(ctx, inv_mass_rho, inv_mass_fluid) &
          bind(C, name='axh_upload_mass')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: inv_mass_rho(*), inv_mass_fluid(*)
     end function

     integer(c_int) function axh_upload_mesh_maps(ctx, igloc_solid, igloc_fluid, &
          axis_solid, axis_fluid) bind(C, name='axh_upload_mesh_maps')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: igloc_solid(*), igloc_fluid(*)
       integer(c_int32_t), intent(in) :: axis_solid(*), axis_fluid(*)
     end function

     integer(c_int) function axh_upload_boundary(ctx, bse, bfe, bjs, bjf, bmatr) &
          bind(C, name='axh_upload_boundary')
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: bse(*), bfe(*), bjs(*), bjf(*)
       real(c_float), intent(in) :: bmatr(*)
     end function

     integer(c_int) function axh_upload_fluid_surface_mask(ctx, mask) &
          bind(C, name='axh_upload_fluid_surface_mask')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: mask(*)
     end function

     integer(c_int) function axh_upload_absorbing(ctx, gsol, gflu) &
          bind(C, name='axh_upload_absorbing')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: gsol(*), gflu(*)
     end function

     integer(c_int) function axh_upload_source(ctx, ielsrc, source_term, stf, nstf) &
          bind(C, name='axh_upload_source')
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: ielsrc(*)
       real(c_float), intent(in) :: source_term(*), stf(*)
       integer(c_int32_t), value :: nstf
     end function

     integer(c_int) function axh_upload_receivers(ctx, recfile_el) &
          bind(C, name='axh_upload_receivers')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: recfile_el(*)
     end function

     integer(c_int) function axh_run(ctx, it_begin, it_end, seis) bind(C, name='axh_run')
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: it_begin, it_end
       real(c_float), intent(out) :: seis(*)
     end function

     integer(c_int) function axh_run_begin(ctx, it_begin, it_end, seis) &
          bind(C, name='axh_run_begin')
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: it_begin, it_end
       real(c_float) :: seis(*)
     end function

     integer(c_int) function axh_run_end(ctx) bind(C, name='axh_run_end')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function

     integer(c_int) function axh_upload_fluid_gradient(ctx, inv_rho, dsdeta, dzdeta, &
          dsdxi, dzdxi) bind(C, name='axh_upload_fluid_gradient')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: inv_rho(*), dsdeta(*), dzdeta(*), dsdxi(*), dzdxi(*)
     end function

     integer(c_int) function axh_upload_kwf_grid(ctx, npoint_solid, npoint_fluid, &
          mapping, mask) bind(C, name='axh_upload_kwf_grid')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: npoint_solid, npoint_fluid
       integer(c_int32_t), intent(in) :: mapping(*), mask(*)
     end function

     integer(c_int) function axh_set_kwf(ctx, strain_it, src_mask, ring_snapshots) &
          bind(C, name='axh_set_kwf')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: strain_it, src_mask, ring_snapshots
     end function

     integer(c_int) function axh_kwf_snapshots(ctx, it_begin, it_end) &
          bind(C, name='axh_kwf_snapshots')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: it_begin, it_end
     end function

     integer(c_int) function axh_run_kwf(ctx, it_begin, it_end, seis, kwf) &
          bind(C, name='axh_run_kwf')
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: it_begin, it_end
       real(c_float), intent(out) :: seis(*), kwf(*)
     end function

     integer(c_int) function axh_run_kwf_begin(ctx, it_begin, it_end, seis, kwf) &
          bind(C, name='axh_run_kwf_begin')
       import :: c_int, c_ptr, c_int32_t, c_float
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: it_begin, it_end
       real(c_float) :: seis(*), kwf(*)
     end function

     integer(c_int) function axh_download_fields(ctx, disp, velo, chi, dchi, ddchi) &
          bind(C, name='axh_download_fields')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(out) :: disp(*), velo(*), chi(*), dchi(*), ddchi(*)
     end function

     integer(c_int) function axh_check_inputs(ctx) bind(C, name='axh_check_inputs')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function

     integer(c_int) function axh_save_state(ctx, path) bind(C, name='axh_save_state')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
     end function

     integer(c_int) function axh_symplectic_coefficients(scheme, deltat, coefd, coefv, &
          coeff, nstages) bind(C, name='axh_symplectic_coefficients')
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: scheme
       real(c_double), value :: deltat
       real(c_double), intent(out) :: coefd(*), coefv(*), coeff(*)
       integer(c_int32_t), intent(out) :: nstages
     end function

     integer(c_int) function axh_convolve_traces(nt, ntraces, seis, nj, stf, dt, scale, &
          out) bind(C, name='axh_convolve_traces')
       import :: c_int, c_int32_t, c_float, c_double
       integer(c_int32_t), value :: nt, ntraces, nj
       real(c_float), intent(in) :: seis(*), stf(*)
       real(c_double), value :: dt, scale
       real(c_float), intent(out) :: out(*)
     end function

     ! host glue (axh_host_glue.c)
     integer(c_int) function axh_glue_reference_sls(nmax, exp_w, ts_t, ts_tm1, y_j, &
          do_corr) bind(C, name='axh_glue_reference_sls')
       import :: c_int, c_double
       integer(c_int), value :: nmax
       real(c_double), intent(out) :: exp_w(*), ts_t(*), ts_tm1(*), y_j(*)
       integer(c_int), intent(out) :: do_corr
     end function

     integer(c_int) function axh_glue_upload_att(ctx, Y, vse, vsx, vze, vzx, dsde, dzde, &
          dsdx, dzdx, invs, dmu, dkap, expw, tst, tstm1, ajmu, ajkap) &
          bind(C, name='axh_glue_upload_att')
       import :: c_int, c_ptr, c_float, c_double
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: Y(*), vse(*), vsx(*), vze(*), vzx(*)
       real(c_float), intent(in) :: dsde(*), dzde(*), dsdx(*), dzdx(*), invs(*)
       real(c_float), intent(in) :: dmu(*), dkap(*)
       real(c_double), intent(in) :: expw(*), tst(*), tstm1(*), ajmu(*), ajkap(*)
     end function

     integer(c_int) function axh_glue_upload_att_full(ctx, Y, vse, vsx, vze, vzx, y0, v0se, &
          v0sx, v0ze, v0zx, dsde, dzde, dsdx, dzdx, invs, dmu, dkap, expw, tst, tstm1, &
          ajmu, ajkap) bind(C, name='axh_glue_upload_att_full')
       import :: c_int, c_ptr, c_float, c_double
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: Y(*), vse(*), vsx(*), vze(*), vzx(*)
       real(c_float), intent(in) :: y0(*), v0se(*), v0sx(*), v0ze(*), v0zx(*)
       real(c_float), intent(in) :: dsde(*), dzde(*), dsdx(*), dzdx(*), invs(*)
       real(c_float), intent(in) :: dmu(*), dkap(*)
       real(c_double), intent(in) :: expw(*), tst(*), tstm1(*), ajmu(*), ajkap(*)
     end function

     integer(c_int) function axh_glue_getenv(name, val, n) bind(C, name='axh_glue_getenv')
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: name(*)
       character(kind=c_char), intent(out) :: val(*)
       integer(c_int), value :: n
     end function

     integer(c_int) function axh_glue_copy_file(from, to) bind(C, name='axh_glue_copy_file')
       import :: c_int, c_char
       character(kind=c_char), intent(in) :: from(*), to(*)
     end function
  end interface

contains

  !> Message of the last library error as a Fortran string.
  function axh_error_string() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    msg = ' '
    cp = axh_last_error()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [512])
    do i = 1, 512
       if (p(i) == c_null_char) exit
       msg(i:i) = p(i)
    enddo
  end function

  !> Stop with the library's message when a call failed (pcheck-style,
  !! SOLVER/commpi.F90:64-111).
  subroutine axh_check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= AXH_OK) then
       write(6,*) 'GPU time loop ERROR in ', trim(what), ': ', trim(axh_error_string())
       call flush(6)
       stop 2
    endif
  end subroutine

end module axh_iface
!=========================================================================================
