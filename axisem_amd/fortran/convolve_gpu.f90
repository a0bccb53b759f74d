!=========================================================================================
!> GPU replacement of convolve_with_stf of the reference's post-processing program
!! (SOLVER/UTILS/post_processing.F90:1014-1084), linked in its place: same interface,
!! same outputs.  The source time function is tabulated once per call in the
!! reference's precisions (its N_j values at tau_j = j dt), the convolution of the
!! three components -- N_j terms for each of the nt samples, the costly part of the
!! post-processing -- runs on the GPU (axh_convolve_traces, the reference's arithmetic
!! bit for bit), and the tabulated function is written to the same file as the
!! reference writes it.
subroutine convolve_with_stf(t_0, dt, nt, stf, outdir, seis, seis_fil)

  use iso_c_binding
  use global_par, only: pi, decay, shift_fact1, dp
  use axh_iface,  only: axh_convolve_traces
  implicit none

  integer, intent(in)            :: nt
  real(kind=dp), intent(in)      :: t_0, dt
  character(len=100), intent(in) :: outdir
  character(len=7), intent(in)   :: stf
  real, intent(in)               :: seis(nt,3)
  real, intent(out)              :: seis_fil(nt,3)

  real, allocatable              :: table(:)
  real                           :: tau, arg, alpha, sqrt_pi_inv, t
  integer                        :: j, nj, u
  integer(c_int)                 :: ierr
  character(len=4)               :: appidur

  ! the number of terms and the Gaussian's constants as the reference forms them
  nj = int(2. * shift_fact1 * t_0 / dt)
  if (nj > nt) nj = nt
  alpha = decay / t_0
  sqrt_pi_inv = 1. / dsqrt(pi)

  allocate(table(max(nj, 1)))
  do j = 1, nj
     tau = dble(j) * dt
     select case (stf)
     case ('gauss_0')
        arg = alpha * (tau - shift_fact1 * t_0)
        if (arg < 50.) then
           table(j) = alpha * exp(-arg**2) * sqrt_pi_inv / pi
        else
           table(j) = 0.
        endif
     case ('gauss_1')
        table(j) = -2. * (decay / t_0)**2 * (tau - shift_fact1 * t_0) * &
                   exp(-((decay / t_0 * (tau - shift_fact1 * t_0))**2))
        table(j) = table(j) / (decay / t_0 * sqrt(2.) * exp(-2.))
     case default
        write(6,*) ' other source time function not implemented yet!', stf
        stop 2
     end select
  enddo

  write(6,*)
  write(6,*) 'Convolving with period=', t_0, ' on the GPU,', nj, 'terms'
  ierr = axh_convolve_traces(int(nt, c_int32_t), 3_c_int32_t, seis, int(nj, c_int32_t), &
                             table, real(dt, c_double), real(pi, c_double), seis_fil)
  if (ierr /= 0) then
     write(6,*) 'GPU convolution failed, status', ierr
     stop 2
  endif

  ! the tabulated source time function, as the reference writes it
  call define_io_appendix(appidur, int(t_0))
  open(newunit=u, file=trim(outdir)//'/stf_'//trim(stf)//'_'//appidur//'sec.dat')
  do j = 1, nj
     t = dt * real(j)
     write(u,*) t, table(j)
  enddo
  close(u)
  deallocate(table)

end subroutine convolve_with_stf
!=========================================================================================
