!=========================================================================================
!> Drop-in replacement of the reference module time_evol_wave
!! (SOLVER/time_evol_wave.F90:26-41): same name, same two public entry points,
!! prepare_waves and time_loop, no arguments.  prepare_waves follows the reference
!! orchestration (SOLVER/time_evol_wave.F90:47-225) by calling the reference's own
!! public setup routines, then hands every array the time loop reads to the GPU
!! library through iso_c_binding (include/axihip.h).  time_loop runs the loop on the
!! GPU and writes the receiver seismograms in the reference's ASCII format
!! (SOLVER/seismograms.f90:742-780) and the generic epicentre/equator/antipode traces
!! (SOLVER/seismograms.f90:688-735).
!!
!! Wavefield snapshots (SAVE_SNAPSHOTS, xdmf and vtk) are written by the reference's own
!! writers from the fields copied back at every snap_it-th step; SAVE_ENERGY copies the
!! fields back after every step (the reference computes the energies every step).
!!
!! Environment switches of the host:
!!   AXH_SAVE_STATE=<path>  write the uploads to a state file (axh_save_state)
!!   AXH_NO_RUN=1           stop after prepare/upload (records states without a GPU)
!!   AXH_EXTRA_SOURCES=<f>  with AXH_NO_RUN: after the first state, record one state per
!!                          line of file f ("<inparam_source file> <state file>"), all
!!                          on the mesh and model terms set up once (one source type);
!!                          f may be a pipe that a request loop writes as requests come
!!   AXH_KWF_STRAIN_IT=<n>  kernel-wavefield dumps (dump_type displ_only) every n steps on
!!                          the grid of the reference's build_kwf_grid: the values the
!!                          NetCDF variables disp_s, disp_p, disp_z of a snapshot hold
!!                          (SOLVER/nc_routines.F90:945-970), appended per snapshot to
!!                          Data/kwf_disp_<s|p|z>.bin (float32, npoint_kwf per snapshot;
!!                          no NetCDF library here, so no axisem_output.nc4)
module time_evol_wave

  use global_parameters
  use data_proc
  use data_mesh
  use data_source
  use data_time
  use seismograms
  use rotations
  use data_io,          only: verbose
  use axh_iface
  ! private helpers of the reference's own time_evol_wave and attenuation modules,
  ! compiled at build time from the reference sources under these module names
  use axh_ref_time_evol_wave, only: create_fluid_free_surface_mask, &
                                    create_absorbing_gamma, ref_energy => energy
  use axh_ref_attenuation,    only: ref_fast_correct => fast_correct


  implicit none
  public :: prepare_waves, time_loop
  private

  type(c_ptr), save :: ctx = c_null_ptr
  integer,     save :: nrec_all = 0          !< receivers + generic traces
  integer,     save :: iepi = 0, ianti = 0, iequ = 0   !< rows of the generic traces
  logical,     save :: no_run = .false.
  integer,     save :: isnap = 0             !< snapshots written (dump_stuff's isnap)
  integer,     save :: kwf_it = 0            !< kernel-wavefield interval (0: no dumps)

contains

!-----------------------------------------------------------------------------------------
!> Reference orchestration of SOLVER/time_evol_wave.F90:47-225, restricted to the outputs
!! the GPU loop produces; requests for other outputs stop with a message.
subroutine prepare_waves

  use data_io
  use parameters
  use def_precomp_terms
  use source
  use clocks_mod
  use commun

  character(len=1024) :: path
  integer(c_int)      :: lpath

  if (lpr) then
     write(6,*)''
     write(6,*)'++++++++++++++++++++++++++++++++++++++++++++++++++++++++++++++++'
     write(6,*)'++++++++   SEISMIC WAVE PROPAGATION ON THE GPU (gfx950)   +++++++'
     write(6,*)'++++++++++++++++++++++++++++++++++++++++++++++++++++++++++++++++'
  endif

  call check_supported_pre()

  ! kernel-wavefield dumps (AXH_KWF_STRAIN_IT) need the fluid displacement, as
  ! readin_parameters decides for dump_type displ_only (SOLVER/parameters.F90:214-221):
  ! read_model_compute_terms then fills inv_rho_fluid
  lpath = axh_glue_getenv('AXH_KWF_STRAIN_IT'//c_null_char, path, len(path))
  if (lpath > 0) need_fluid_displ = .true.

  call read_sourceparams
  if (rot_src) call def_rot_matrix
  call compute_numerical_parameters
  call read_model_compute_terms
  call compute_stf
  call compute_src
  call create_fluid_free_surface_mask()
  call create_absorbing_gamma()

  ! energy files of prepare_waves (SOLVER/time_evol_wave.F90:102-110)
  if (dump_energy .and. lpr) then
     write(6,*)'  opening files for kinetic/potential energy...'
     if (have_fluid) then
        open(unit=4444, file=datapath(1:lfdata)//'/energy_sol.dat')
        open(unit=4445, file=datapath(1:lfdata)//'/energy_flu.dat')
     endif
     open(unit=4446, file=datapath(1:lfdata)//'/energy_glob.dat')
  endif

  call prepare_snapshots()

  call prepare_seismograms
  call open_hyp_epi_equ_anti
  call prepare_from_recfile_seis

  call write_parameters
  call upload_all()

  lpath = axh_glue_getenv('AXH_SAVE_STATE'//c_null_char, path, len(path))
  if (lpath > 0) then
     call axh_check(axh_save_state(ctx, trim(path(1:lpath))//c_null_char), 'save state')
     if (lpr) write(6,'(a,a)') ' wrote GPU state file ', trim(path(1:lpath))
     call flush(6)
  endif
  lpath = axh_glue_getenv('AXH_NO_RUN'//c_null_char, path, len(path))
  no_run = lpath > 0 .and. path(1:1) /= '0'
  if (no_run) call record_extra_sources()

  call barrier
  if (lpr) write(6,*) 'done preparing waves (GPU time loop).'

end subroutine prepare_waves
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> Entry point of the loop (SOLVER/time_evol_wave.F90:231-245): runs all niter steps on
!! the GPU in chunks and writes the seismogram samples after each chunk.
subroutine time_loop

  use clocks_mod, only: tick
  use data_io,    only: datapath, lfdata

  integer                          :: it0, it1, nsamp, k, ierr, chunk, isamp_iter
  integer                          :: nk, ncomp, npt, c
  real(kind=realkind), allocatable :: seis(:), kwf(:)
  real(kind=dp)                    :: tloc
  character(len=1)                 :: kname(3)

  if (no_run) then
     if (lpr) write(6,*) 'AXH_NO_RUN set: skipping the time loop.'
     return
  endif

  iclockold = tick()

  if (lpr) write(6,*) '************ S T A R T I N G   T I M E   L O O P *************'

  ! dump_stuff(0, ...) of the reference samples the zero initial state at t = 0
  ! (SOLVER/time_evol_wave.F90:347,579)
  if (nrec_all > 0) then
     allocate(seis(3 * nrec_all))
     seis = 0
     call write_samples(seis, 1, 0.0_dp)
     deallocate(seis)
  endif

  ! dump_stuff(0, ...) also writes the snapshot and the energies of the initial state
  if (dump_vtk .or. dump_xdmf) call write_snapshot(0.0_dp)
  if (dump_energy) call energy_gpu(0.0_dp)

  ! chunks of at most 512 seismogram samples, ending at the snapshot steps (every
  ! step with SAVE_ENERGY)
  chunk = max(1, min(niter, 512 * seis_it))
  if (dump_energy) chunk = 1
  if (kwf_it > 0) then
     ! kernel wavefields: chunks of at most 16 snapshots, one file per variable
     chunk = min(chunk, 16 * kwf_it)
     npt = npoint_solid_kwf + npoint_fluid_kwf
     if (src_type(1) == 'monopole') then
        ncomp = 2
        kname(1:2) = ['s', 'z']
     else
        ncomp = 3
        kname = ['s', 'p', 'z']
     endif
     allocate(kwf(max(1, (chunk / kwf_it + 2) * ncomp * npt)))
     do c = 1, ncomp
        open(13200 + c, file=datapath(1:lfdata)//'/kwf_disp_'//kname(c)//'.bin', &
             access='stream', form='unformatted', status='replace')
     enddo
  endif
  allocate(seis(3 * max(nrec_all, 1) * (chunk / seis_it + 1)))
  tloc = zero
  isamp_iter = 0
  it0 = 1
  if (.not. (dump_vtk .or. dump_xdmf .or. dump_energy)) then
     ! nothing needs the fields between chunks: the next chunk runs on the GPU
     ! while this one's samples (and kernel-wavefield snapshots) are written
     if (kwf_it == 0) then
        ncomp = 0
        npt = 0
        allocate(kwf(1))
     endif
     call time_loop_overlapped(seis, kwf, chunk, tloc, ncomp, npt)
     it0 = niter + 1
  endif
  do while (it0 <= niter)
     it1 = min(niter, it0 + chunk - 1)
     if (dump_vtk .or. dump_xdmf) it1 = min(it1, (it0 + snap_it - 1) / snap_it * snap_it)
     if (kwf_it > 0) then
        nk = axh_kwf_snapshots(ctx, int(it0, c_int32_t), int(it1, c_int32_t))
        ierr = axh_run_kwf(ctx, int(it0, c_int32_t), int(it1, c_int32_t), seis, kwf)
     else
        ierr = axh_run(ctx, int(it0, c_int32_t), int(it1, c_int32_t), seis)
     endif
     if (ierr == AXH_ERR_BLOWUP) then
        write(6,*) procstrg, '!!!!!!!!!!!!!!! DISPLACEMENTS BLEW UP !!!!!!!!!!!!!!!'
        write(6,*) procstrg, trim(axh_error_string())
        stop
     endif
     call axh_check(ierr, 'axh_run')
     ! the chunk's kernel-wavefield snapshots, variable by variable
     if (kwf_it > 0) then
        do k = 1, nk
           do c = 1, ncomp
              write(13200 + c) kwf(1 + ((k - 1) * ncomp + c - 1) * npt:((k - 1) * ncomp + c) * npt)
           enddo
        enddo
     endif
     ! replay the reference's time accumulation t = t + deltat for the sampled steps
     nsamp = 0
     do k = it0, it1
        tloc = tloc + deltat
        if (mod(k, seis_it) == 0 .and. nrec_all > 0) then
           nsamp = nsamp + 1
           call write_samples(seis(1 + 3 * nrec_all * (nsamp - 1):), 1, tloc)
        endif
     enddo
     ! (nested: Fortran need not short-circuit, and snap_it is 0 without snapshots)
     if (dump_vtk .or. dump_xdmf) then
        if (mod(it1, snap_it) == 0) call write_snapshot(tloc)
     endif
     if (dump_energy) call energy_gpu(tloc)
     if (lpr) then
        write(6,13) it1, tloc, real(it1) / real(niter) * 100.
        call flush(6)
     endif
13   format('  time step:',i6,'; t=',f8.2,' s (',f5.1,'%)')
     it0 = it1 + 1
  enddo
  t = tloc
  if (kwf_it > 0) then
     do c = 1, ncomp
        close(13200 + c)
     enddo
  endif

  iclockold = tick(id=idold, since=iclockold)

  ierr = axh_destroy(ctx)
  ctx = c_null_ptr

end subroutine time_loop
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> Snapshot grids and files of prepare_waves (SOLVER/time_evol_wave.F90:117-175): the
!! vtk midpoint grid, and the xdmf grid with its stream files (big endian, appended).
subroutine prepare_snapshots()

  use data_io
  use meshes_io, only : dump_glob_grid_midpoint, dump_xdmf_grid
  character(len=512) :: fname

  if (dump_vtk) then
     if (lpr) write(6,*)'  dumping global grids for snapshots...'
     call dump_glob_grid_midpoint(ibeg, iend, jbeg, jend)
  endif
  if (dump_xdmf) then
     if (lpr) write(6,*)'  dumping mesh for xdmf snapshots...'
     fname = datapath(1:lfdata)//'/xdmf_snap_s_' //appmynum//'.dat'
     open(13100, file=trim(fname), access='stream', status='unknown', &
          form='unformatted', convert='big_endian', position='append')
     if (.not. src_type(1) == 'monopole') then
        fname = datapath(1:lfdata)//'/xdmf_snap_p_' //appmynum//'.dat'
        open(13101, file=trim(fname), access='stream', status='unknown', &
             form='unformatted', convert='big_endian', position='append')
     endif
     fname = datapath(1:lfdata)//'/xdmf_snap_z_' //appmynum//'.dat'
     open(13102, file=trim(fname), access='stream', status='unknown', &
          form='unformatted', convert='big_endian', position='append')
     fname = datapath(1:lfdata)//'/xdmf_snap_trace_' //appmynum//'.dat'
     open(13103, file=trim(fname), access='stream', status='unknown', &
          form='unformatted', convert='big_endian', position='append')
     fname = datapath(1:lfdata)//'/xdmf_snap_curlip_' //appmynum//'.dat'
     open(13104, file=trim(fname), access='stream', status='unknown', &
          form='unformatted', convert='big_endian', position='append')
     call dump_xdmf_grid()
  endif

end subroutine prepare_snapshots
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> The snapshot part of dump_stuff (SOLVER/time_evol_wave.F90:1152-1177) for the
!! current fields, copied back from the GPU (all zero before the first step).
subroutine write_snapshot(time)

  use data_io
  use wavefields_io, only : glob_snapshot_midpoint, glob_snapshot_xdmf
  real(kind=dp), intent(in)        :: time
  real(kind=realkind), allocatable :: disp(:,:,:,:), velo(:,:,:,:), chi(:,:,:)
  real(kind=realkind), allocatable :: dchi(:,:,:), ddchi(:,:,:)

  allocate(disp(0:npol, 0:npol, nel_solid, 3), velo(0:npol, 0:npol, nel_solid, 3))
  allocate(chi(0:npol, 0:npol, max(nel_fluid, 1)), dchi(0:npol, 0:npol, max(nel_fluid, 1)))
  allocate(ddchi(0:npol, 0:npol, max(nel_fluid, 1)))
  call axh_check(axh_download_fields(ctx, disp, velo, chi, dchi, ddchi), 'download fields')
  if (dump_vtk) then
     isnap = isnap + 1
     if (lpr) then
        write(6,*)
        write(6,*) 'Writing global snap to file: ', isnap
        write(6,*)
     endif
     call glob_snapshot_midpoint(disp, chi(:,:,1:nel_fluid), ibeg, iend, jbeg, jend, isnap)
  endif
  if (dump_xdmf) then
     if (.not. dump_vtk) isnap = isnap + 1
     if (lpr) then
        write(6,*)
        write(6,*)'Writing global xdmf snap to file:',isnap
        write(6,*)
     endif
     call glob_snapshot_xdmf(disp, chi(:,:,1:nel_fluid), time, isnap)
  endif
  deallocate(disp, velo, chi, dchi, ddchi)

end subroutine write_snapshot
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> The energies after the step ending at time: the fields copied back from the GPU handed
!! to the reference's own energy routine (SOLVER/time_evol_wave.F90:1424-1526, private
!! there, reached through its build-time module axh_ref_time_evol_wave).
subroutine energy_gpu(time)

  real(kind=dp), intent(in)        :: time
  real(kind=realkind), allocatable :: disp(:,:,:,:), vel(:,:,:,:), chi(:,:,:)
  real(kind=realkind), allocatable :: dchi(:,:,:), ddchi(:,:,:)

  allocate(disp(0:npol, 0:npol, nel_solid, 3), vel(0:npol, 0:npol, nel_solid, 3))
  allocate(chi(0:npol, 0:npol, max(nel_fluid, 1)), dchi(0:npol, 0:npol, max(nel_fluid, 1)))
  allocate(ddchi(0:npol, 0:npol, max(nel_fluid, 1)))
  call axh_check(axh_download_fields(ctx, disp, vel, chi, dchi, ddchi), 'download fields')
  t = time
  call ref_energy(disp, vel, dchi(:,:,1:nel_fluid), ddchi(:,:,1:nel_fluid))
  deallocate(disp, vel, chi, dchi, ddchi)

end subroutine energy_gpu
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> One sample of compute_recfile_seis_bare (SOLVER/seismograms.f90:742-780) and of
!! compute_hyp_epi_equ_anti (SOLVER/seismograms.f90:688-735), from the library's
!! receiver rows s(3, nrec_all) in the same component order.
subroutine write_samples(s, nsamp, time)

  real(kind=realkind), intent(in) :: s(*)
  integer, intent(in)             :: nsamp
  real(kind=dp), intent(in)       :: time
  integer                         :: i, b

  if (nsamp < 1) return
  do i = 1, num_rec
     b = 3 * (i - 1)
     if (src_type(1) == 'monopole') then
        write(100000+i,*) s(b + 1), s(b + 2)
     else
        write(100000+i,*) s(b + 1), s(b + 2), s(b + 3)
     endif
  enddo

  if (iepi > 0) then
     b = 3 * (iepi - 1)
     write(900,*) time, s(b + 1)
     if (src_type(1) /= 'monopole') then
        write(903,*) time, s(b + 2)
        write(906,*) time, s(b + 3)
     else
        write(906,*) time, s(b + 2)
     endif
  endif
  if (ianti > 0) then
     b = 3 * (ianti - 1)
     write(901,*) time, s(b + 1)
     if (src_type(1) /= 'monopole') then
        write(905,*) time, s(b + 2)
        write(908,*) time, s(b + 3)
     else
        write(908,*) time, s(b + 2)
     endif
  endif
  if (iequ > 0) then
     b = 3 * (iequ - 1)
     write(902,*) time, s(b + 1)
     if (src_type(1) /= 'monopole') then
        write(904,*) time, s(b + 2)
        write(907,*) time, s(b + 3)
     else
        write(907,*) time, s(b + 2)
     endif
  endif

end subroutine write_samples
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> The chunks of the time loop with two sample buffers: chunk k + 1 is queued on the GPU
!! (axh_run_begin) before the samples of chunk k are written, and collected
!! (axh_run_end) after, so the writes of the reference's seismogram files overlap the
!! time steps.  Same files as one chunk after the other.
subroutine time_loop_overlapped(seis, kwf, chunk, tloc, ncomp, npt)

  real(kind=realkind), allocatable, intent(inout) :: seis(:), kwf(:)
  integer, intent(in)                             :: chunk, ncomp, npt
  real(kind=dp), intent(inout)                    :: tloc

  real(kind=realkind), allocatable :: next(:), tmp(:), knext(:)
  integer                          :: it0, it1, n0, n1, k, nsamp, c
  integer(c_int)                   :: ierr

  allocate(next(size(seis)), knext(size(kwf)))
  it0 = 1
  it1 = min(niter, chunk)
  call queue_chunk(it0, it1, seis, kwf)
  do
     ierr = axh_run_end(ctx)
     if (ierr == AXH_ERR_BLOWUP) then
        write(6,*) procstrg, '!!!!!!!!!!!!!!! DISPLACEMENTS BLEW UP !!!!!!!!!!!!!!!'
        write(6,*) procstrg, trim(axh_error_string())
        stop
     endif
     call axh_check(ierr, 'axh_run_end')
     n0 = it1 + 1
     n1 = min(niter, it1 + chunk)
     if (n0 <= niter) call queue_chunk(n0, n1, next, knext)
     ! the kernel-wavefield snapshots of steps it0..it1, variable by variable
     if (kwf_it > 0) then
        do k = 1, axh_kwf_snapshots(ctx, int(it0, c_int32_t), int(it1, c_int32_t))
           do c = 1, ncomp
              write(13200 + c) kwf(1 + ((k - 1) * ncomp + c - 1) * npt:((k - 1) * ncomp + c) * npt)
           enddo
        enddo
     endif
     ! the samples of steps it0..it1, with the reference's time accumulation
     nsamp = 0
     do k = it0, it1
        tloc = tloc + deltat
        if (mod(k, seis_it) == 0 .and. nrec_all > 0) then
           nsamp = nsamp + 1
           call write_samples(seis(1 + 3 * nrec_all * (nsamp - 1):), 1, tloc)
        endif
     enddo
     if (lpr) then
        write(6,13) it1, tloc, real(it1) / real(niter) * 100.
        call flush(6)
     endif
13   format('  time step:',i6,'; t=',f8.2,' s (',f5.1,'%)')
     if (n0 > niter) exit
     call move_alloc(seis, tmp)
     call move_alloc(next, seis)
     call move_alloc(tmp, next)
     call move_alloc(kwf, tmp)
     call move_alloc(knext, kwf)
     call move_alloc(tmp, knext)
     it0 = n0
     it1 = n1
  enddo
  deallocate(next, knext)

contains

  subroutine queue_chunk(a, b, s, kb)
    integer, intent(in)                :: a, b
    real(kind=realkind), intent(inout) :: s(:), kb(:)
    if (kwf_it > 0) then
       call axh_check(axh_run_kwf_begin(ctx, int(a, c_int32_t), int(b, c_int32_t), s, kb), &
                      'axh_run_kwf_begin')
    else
       call axh_check(axh_run_begin(ctx, int(a, c_int32_t), int(b, c_int32_t), s), &
                      'axh_run_begin')
    endif
  end subroutine queue_chunk

end subroutine time_loop_overlapped
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> Outputs and options the GPU loop does not produce stop here, before any work.
subroutine check_supported_pre()

  use data_io
  character(len=256) :: why

  why = ' '
  if (nproc > 1)        why = 'more than one mesh slice (NTHETA/NRADIAL_SLICES > 1)'
  if (npol /= 4)        why = 'npol /= 4'
  if (dump_wavefields)  why = 'KERNEL_WAVEFIELDS (NetCDF; AXH_KWF_STRAIN_IT writes displ_only)'
  if (use_netcdf)       why = 'USE_NETCDF (ASCII seismograms only)'
  if (why /= ' ') then
     write(6,*) 'GPU time loop: unsupported option: ', trim(why)
     stop 2
  endif

end subroutine check_supported_pre
!-----------------------------------------------------------------------------------------


!-----------------------------------------------------------------------------------------
!> Hands every array of the loop to the library (and optionally to a state file).
subroutine upload_all()

  use data_matr
  use data_spec
  use data_pointwise
  use data_io,        only: datapath, lfdata, need_fluid_displ
  use attenuation,    only: n_sls_attenuation, att_coarse_grained
  use source,         only: compute_stf_t

  type(axh_dims_t)              :: dims
  integer(c_int)                :: ierr
  integer(c_int32_t), allocatable :: ax_s(:), ax_f(:), rec(:,:), isrc(:)
  real(kind=realkind), allocatable :: stf_tab(:), empty(:)
  real(kind=dp)                 :: coefd(8), coefv(8), coeff(8), subdt(8), stf_symp(8)
  real(kind=dp)                 :: tacc
  integer(c_int32_t)            :: nst
  integer                       :: i, iter, nextra, iel

  ! time scheme (read_inparam_basic lower-cases TIME_SCHEME,
  ! SOLVER/parameters.F90:471, so the McLachlan schemes arrive as ml_so*)
  select case (trim(time_scheme))
  case ('newmark2');  dims%scheme = AXH_NEWMARK2
  case ('symplec4');  dims%scheme = AXH_SYMPLEC4
  case ('ML_SO4m5', 'ml_so4m5');  dims%scheme = AXH_ML_SO4M5
  case ('ML_SO6m7', 'ml_so6m7');  dims%scheme = AXH_ML_SO6M7
  case default
     write(6,*) 'GPU time loop: unsupported time scheme ', trim(time_scheme)
     stop 2
  end select

  select case (trim(src_type(1)))
  case ('monopole'); dims%src_type = AXH_MONOPOLE
  case ('dipole');   dims%src_type = AXH_DIPOLE
  case ('quadpole'); dims%src_type = AXH_QUADPOLE
  case default
     write(6,*) 'GPU time loop: unknown source type ', trim(src_type(1))
     stop 2
  end select

  dims%abi_version = AXH_ABI_VERSION
  dims%npol        = npol
  dims%nel_solid   = nel_solid
  dims%nel_fluid   = nel_fluid
  dims%nglob_solid = nglob_solid
  dims%nglob_fluid = nglob_fluid
  dims%nel_bdry    = nel_bdry
  dims%anel        = merge(1, 0, anel_true)
  dims%coarse_grained = 0
  dims%n_sls       = 0
  if (anel_true) then
     dims%coarse_grained = merge(1, 0, att_coarse_grained)
     dims%n_sls = n_sls_attenuation
  endif
  dims%niter     = niter
  dims%seis_it   = seis_it
  dims%nelsrc    = nelsrc
  dims%fluid_src = merge(1, 0, fluid_src)
  dims%reserved  = 0
  dims%deltat    = deltat
  dims%magnitude = magnitude

  ! receivers, then the generic epicentre / antipode / equator traces
  nextra = 0
  if (maxind > 0) then
     if (have_epi)      nextra = nextra + 1
     if (have_antipode) nextra = nextra + 1
     if (have_equ)      nextra = nextra + 1
  endif
  nrec_all = num_rec + nextra
  dims%num_rec = nrec_all
  allocate(rec(max(nrec_all,1), 3))
  do i = 1, num_rec
     rec(i, :) = recfile_el(i, :)
  enddo
  i = num_rec
  if (maxind > 0) then
     if (have_epi) then
        i = i + 1; iepi = i;  rec(i, :) = [ielepi, 0, npol]
     endif
     if (have_antipode) then
        i = i + 1; ianti = i; rec(i, :) = [ielantipode, 0, 0]
     endif
     if (have_equ) then
        i = i + 1; iequ = i;  rec(i, :) = [ielequ, npol, npol]
     endif
  endif

  call axh_check(axh_create(dims, ctx), 'axh_create')

  call axh_check(axh_upload_spectral(ctx, G0, G1, G1T, G2, G2T), 'spectral')

  call up3d(AXH_M11S, M11s); call up3d(AXH_M21S, M21s); call up3d(AXH_M41S, M41s)
  call up3d(AXH_M12S, M12s); call up3d(AXH_M22S, M22s); call up3d(AXH_M42S, M42s)
  call up3d(AXH_M11Z, M11z); call up3d(AXH_M21Z, M21z); call up3d(AXH_M41Z, M41z)
  call up3d(AXH_M32S, M32s); call up3d(AXH_M13S, M13s); call up3d(AXH_M33S, M33s)
  call up3d(AXH_M43S, M43s); call up3d(AXH_M1PHI, M1phi); call up3d(AXH_M2PHI, M2phi)
  call up3d(AXH_M4PHI, M4phi)
  call up3d(AXH_M_1, M_1); call up3d(AXH_M_2, M_2); call up3d(AXH_M_3, M_3)
  call up3d(AXH_M_4, M_4); call up3d(AXH_M_5, M_5); call up3d(AXH_M_6, M_6)
  call up3d(AXH_M_7, M_7); call up3d(AXH_M_8, M_8)
  call up3d(AXH_M_W1, M_w1); call up3d(AXH_M_W2, M_w2); call up3d(AXH_M_W3, M_w3)
  call up3d(AXH_M_W4, M_w4); call up3d(AXH_M_W5, M_w5)
  call upax(AXH_M0_W1 + 0, M0_w1); call upax(AXH_M0_W1 + 1, M0_w2)
  call upax(AXH_M0_W1 + 2, M0_w3); call upax(AXH_M0_W1 + 3, M0_w4)
  call upax(AXH_M0_W1 + 4, M0_w5); call upax(AXH_M0_W1 + 5, M0_w6)
  call upax(AXH_M0_W1 + 6, M0_w7); call upax(AXH_M0_W1 + 7, M0_w8)
  call upax(AXH_M0_W1 + 8, M0_w9); call upax(AXH_M0_W1 + 9, M0_w10)

  allocate(empty(1))
  empty = 0
  if (nel_fluid > 0) then
     if (allocated(M1chi_fl)) &
        call axh_check(axh_upload_fluid_term(ctx, AXH_M1CHI_FL, M1chi_fl), 'M1chi_fl')
     if (allocated(M2chi_fl)) &
        call axh_check(axh_upload_fluid_term(ctx, AXH_M2CHI_FL, M2chi_fl), 'M2chi_fl')
     if (allocated(M4chi_fl)) &
        call axh_check(axh_upload_fluid_term(ctx, AXH_M4CHI_FL, M4chi_fl), 'M4chi_fl')
     if (allocated(M_w_fl) .and. src_type(1) /= 'monopole') &
        call axh_check(axh_upload_fluid_term(ctx, AXH_M_W_FL, M_w_fl), 'M_w_fl')
     if (allocated(M0_w_fl) .and. src_type(1) /= 'monopole') &
        call axh_check(axh_upload_fluid_term(ctx, AXH_M0_W_FL, M0_w_fl), 'M0_w_fl')
     call axh_check(axh_upload_mass(ctx, inv_mass_rho, inv_mass_fluid), 'mass')
     call axh_check(axh_upload_fluid_surface_mask(ctx, fluid_free_surface_mask), 'fsmask')
     if (have_absorbing_bc) &
        call axh_check(axh_upload_absorbing(ctx, solid_absorbing_gamma, &
                                            fluid_absorbing_gamma), 'absorbing')
  else
     if (have_absorbing_bc) &
        call axh_check(axh_upload_absorbing(ctx, solid_absorbing_gamma, empty), 'absorbing')
     call axh_check(axh_upload_mass(ctx, inv_mass_rho, empty), 'mass')
  endif

  allocate(ax_s(nel_solid), ax_f(max(nel_fluid,1)))
  ax_s = 0
  ax_f = 0
  do iel = 1, nel_solid
     if (axis_solid(iel)) ax_s(iel) = 1
  enddo
  do iel = 1, nel_fluid
     if (axis_fluid(iel)) ax_f(iel) = 1
  enddo
  if (nel_fluid > 0) then
     call axh_check(axh_upload_mesh_maps(ctx, igloc_solid, igloc_fluid, ax_s, ax_f), 'maps')
  else
     call axh_check(axh_upload_mesh_maps(ctx, igloc_solid, ax_f, ax_s, ax_f), 'maps')
  endif

  if (nel_bdry > 0) &
     call axh_check(axh_upload_boundary(ctx, bdry_solid_el, bdry_fluid_el, &
                    bdry_jpol_solid, bdry_jpol_fluid, bdry_matr), 'boundary')

  ! source time function table
  if (dims%scheme == AXH_NEWMARK2) then
     allocate(stf_tab(niter))
     stf_tab = stf(1:niter)
     nst = 1
  else
     call axh_check(axh_symplectic_coefficients(int(dims%scheme, c_int32_t), deltat, &
                    coefd, coefv, coeff, nst), 'symplectic coefficients')
     allocate(stf_tab(niter * nst))
     ! t accumulation and sub-step times of SOLVER/time_evol_wave.F90:586-593
     tacc = zero
     do iter = 1, niter
        tacc = tacc + deltat
        subdt(1:nst) = tacc - deltat + coeff(1:nst)
        call compute_stf_t(int(nst), subdt(1:nst), stf_symp(1:nst))
        do i = 1, nst
           stf_tab((iter - 1) * nst + i) = real(stf_symp(i), kind=realkind)
        enddo
     enddo
  endif
  allocate(isrc(max(nelsrc,1)))
  isrc = 0
  isrc(1:nelsrc) = ielsrc(1:nelsrc)
  if (fluid_src) then
     call axh_check(axh_upload_source(ctx, isrc, source_term_fl, stf_tab, &
                    int(size(stf_tab), c_int32_t)), 'source')
  else
     call axh_check(axh_upload_source(ctx, isrc, source_term_el, stf_tab, &
                    int(size(stf_tab), c_int32_t)), 'source')
  endif

  if (anel_true) call upload_attenuation()

  if (nrec_all > 0) call axh_check(axh_upload_receivers(ctx, rec), 'receivers')

  ! the fluid terms of the kernel-wavefield dumps (dump_disp_global), where the
  ! reference computes inv_rho_fluid (SOLVER/def_precomp_terms.f90:803-805)
  if (have_fluid .and. need_fluid_displ .and. allocated(inv_rho_fluid) .and. &
      allocated(DsDeta_over_J_flu)) &
     call axh_check(axh_upload_fluid_gradient(ctx, inv_rho_fluid, DsDeta_over_J_flu, &
                    DzDeta_over_J_flu, DsDxi_over_J_flu, DzDxi_over_J_flu), 'fluid gradient')
  call upload_kwf()

  ! layout checks of the library, before any device work
  call axh_check(axh_check_inputs(ctx), 'input checks')


contains

  subroutine up3d(term, a)
    integer(c_int), intent(in) :: term
    real(kind=realkind), allocatable, intent(in) :: a(:,:,:)
    if (allocated(a)) call axh_check(axh_upload_solid_term(ctx, term, a), 'solid term')
  end subroutine

  subroutine upax(term, a)
    integer(c_int), intent(in) :: term
    real(kind=realkind), allocatable, intent(in) :: a(:,:)
    if (allocated(a)) call axh_check(axh_upload_solid_term(ctx, term, a), 'axial term')
  end subroutine

end subroutine upload_all
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> With AXH_KWF_STRAIN_IT set: the kernel-wavefield grid of the reference's own
!! build_kwf_grid (SOLVER/meshes_io.F90:489-778; ibeg..jend and the KERNEL_* limits of
!! inparam_advanced) and the dump interval, handed to the library.
subroutine upload_kwf()

  use meshes_io, only: build_kwf_grid

  character(len=64)               :: val
  integer(c_int)                  :: lval
  integer(c_int32_t), allocatable :: kmask(:,:,:)
  integer                         :: ios

  lval = axh_glue_getenv('AXH_KWF_STRAIN_IT'//c_null_char, val, len(val))
  if (lval <= 0) return
  read(val(1:lval), *, iostat=ios) kwf_it
  if (ios /= 0 .or. kwf_it < 1) then
     write(6,*) 'AXH_KWF_STRAIN_IT must be a positive step count: ', val(1:lval)
     stop 2
  endif
  call build_kwf_grid()
  allocate(kmask(0:npol, 0:npol, nel_solid + nel_fluid))
  kmask = merge(1_c_int32_t, 0_c_int32_t, kwf_mask(:, :, 1:nel_solid + nel_fluid))
  call axh_check(axh_upload_kwf_grid(ctx, int(npoint_solid_kwf, c_int32_t), &
                 int(npoint_fluid_kwf, c_int32_t), mapping_ijel_ikwf, kmask), 'kwf grid')
  call axh_check(axh_set_kwf(ctx, int(kwf_it, c_int32_t), &
                 merge(1_c_int32_t, 0_c_int32_t, trim(src_dump_type) == 'mask'), 0_c_int32_t), &
                 'kwf')
  if (lpr) write(6,'(a,i9,a,i9,a,i6,a)') '  kernel wavefields: ', npoint_solid_kwf, &
     ' solid +', npoint_fluid_kwf, ' fluid points every', kwf_it, ' steps'

end subroutine upload_kwf
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> Record mode with several sources of one source type (AXH_EXTRA_SOURCES): the mesh,
!! the model and every precomputed term stay as set up for the first source; for each
!! listed source the reference's own read_sourceparams, compute_stf and compute_src run
!! again (on a copy of the listed file as ./inparam_source, where read_sourceparams
!! reads it, SOLVER/source.f90:51) and the uploads are written to the listed state file.
!! Each state equals the one a separate run of the host on that source records.
subroutine record_extra_sources()

  use source, only: read_sourceparams, compute_stf, compute_src

  character(len=1024) :: list, line, srcfile, statefile
  character(len=10)   :: type0
  integer(c_int)      :: l
  integer             :: u, ios, nrec, i

  l = axh_glue_getenv('AXH_EXTRA_SOURCES'//c_null_char, list, len(list))
  if (l <= 0) return
  type0 = src_type(1)
  nrec = nrec_all
  open(newunit=u, file=list(1:l), status='old', action='read', iostat=ios)
  if (ios /= 0) then
     write(6,*) 'GPU time loop: cannot open the source list ', list(1:l)
     stop 2
  endif
  do
     read(u, '(a)', iostat=ios) line
     if (ios /= 0) exit
     if (len_trim(line) == 0) cycle
     ! two blank-separated paths (not list-directed input: '/' would end the record)
     line = adjustl(line)
     i = index(trim(line), ' ')
     if (i == 0) then
        write(6,*) 'GPU time loop: bad line in the source list: ', trim(line)
        stop 2
     endif
     srcfile = line(1:i-1)
     statefile = adjustl(line(i+1:))
     if (axh_glue_copy_file(trim(srcfile)//c_null_char, 'inparam_source'//c_null_char) /= 0) then
        write(6,*) 'GPU time loop: cannot copy ', trim(srcfile), ' to inparam_source'
        stop 2
     endif
     call read_sourceparams
     if (src_type(1) /= type0 .or. rot_src) then
        write(6,*) 'GPU time loop: the sources of one record run must be on the axis and ', &
                   'of one source type (', trim(type0), '); ', trim(srcfile), ' is ', &
                   trim(src_type(1))
        stop 2
     endif
     if (allocated(stf)) deallocate(stf)
     if (allocated(source_term_el)) deallocate(source_term_el)
     if (allocated(source_term_fl)) deallocate(source_term_fl)
     ielsrc = 0
     call compute_stf
     call compute_src
     call axh_check(axh_destroy(ctx), 'destroy')
     ctx = c_null_ptr
     call upload_all()
     if (nrec_all /= nrec) then
        write(6,*) 'GPU time loop: receivers changed between sources'
        stop 2
     endif
     call axh_check(axh_save_state(ctx, trim(statefile)//c_null_char), 'save state')
     ! a reader waiting for this state (a pipe of requests) sees it at once
     if (lpr) write(6,'(a,a)') ' wrote GPU state file ', trim(statefile)
     call flush(6)
  enddo
  close(u)

end subroutine record_extra_sources
!-----------------------------------------------------------------------------------------

!-----------------------------------------------------------------------------------------
!> Attenuation terms, coarse-grained or at all 25 points (att_coarse_grained).  The
!! per-element SLS weights a_j are those of time_step_memvars_cg4/_4
!! (SOLVER/attenuation.f90:105-127, 244-262), evaluated for every element from the
!! reference's own fitted y_j and Q values.
subroutine upload_attenuation()

  use data_matr
  use data_pointwise
  use attenuation,  only: n_sls_attenuation, att_coarse_grained

  real(kind=dp)       :: expw(64), tst(64), tstm1(64), yj(64)
  integer(c_int)      :: docorr, n
  real(kind=dp), allocatable :: ajmu(:,:), ajk(:,:), yp(:)
  real(kind=realkind), allocatable :: invs(:,:)
  integer             :: iel

  n = axh_glue_reference_sls(64, expw, tst, tstm1, yj, docorr)
  if (n /= n_sls_attenuation) then
     write(6,*) 'GPU time loop: could not read the SLS coefficients of the attenuation module'
     stop 2
  endif
  allocate(ajmu(n, nel_solid), ajk(n, nel_solid), yp(n))
  do iel = 1, nel_solid
     if (docorr /= 0) then
        call ref_fast_correct(yj(1:n) / Q_mu(iel), yp)
     else
        yp = yj(1:n) / Q_mu(iel)
     endif
     ajmu(:, iel) = yp / sum(yp)
     if (docorr /= 0) then
        call ref_fast_correct(yj(1:n) / Q_kappa(iel), yp)
     else
        yp = yj(1:n) / Q_kappa(iel)
     endif
     ajk(:, iel) = yp / sum(yp)
  enddo

  if (.not. att_coarse_grained) then
     call axh_check(axh_glue_upload_att_full(ctx, Y, V_s_eta, V_s_xi, V_z_eta, V_z_xi, &
                    Y0, V0_s_eta, V0_s_xi, V0_z_eta, V0_z_xi, DsDeta_over_J_sol, &
                    DzDeta_over_J_sol, DsDxi_over_J_sol, DzDxi_over_J_sol, inv_s_solid, &
                    delta_mu, delta_kappa, expw, tst, tstm1, ajmu, ajk), 'attenuation')
     return
  endif

  ! f_over_s_solid_el_cg4 reads inv_s_solid at (1,1),(1,3),(3,1),(3,3)
  ! (SOLVER/pointwise_derivatives.f90:66-69)
  allocate(invs(4, nel_solid))
  invs(1,:) = inv_s_solid(1,1,:)
  invs(2,:) = inv_s_solid(1,3,:)
  invs(3,:) = inv_s_solid(3,1,:)
  invs(4,:) = inv_s_solid(3,3,:)

  call axh_check(axh_glue_upload_att(ctx, Y_cg4, V_s_eta_cg4, V_s_xi_cg4, V_z_eta_cg4, &
                 V_z_xi_cg4, DsDeta_over_J_sol_cg4, DzDeta_over_J_sol_cg4, &
                 DsDxi_over_J_sol_cg4, DzDxi_over_J_sol_cg4, invs, delta_mu_cg4, &
                 delta_kappa_cg4, expw, tst, tstm1, ajmu, ajk), 'attenuation')

end subroutine upload_attenuation
!-----------------------------------------------------------------------------------------

end module time_evol_wave
!=========================================================================================
