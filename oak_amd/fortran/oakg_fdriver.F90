! Fortran driver of locAnalysis_gpu: a small 1-D local analysis with a
! Gaussian taper callback (OAK's selectObservations interface, called with the
! state index of a zone's first element).  Runs it with localise_obs = .true.
! and .false. and writes the inputs, the outputs and the return codes to a
! binary file for the tests.
!
!   oakg_fdriver <output file>
module fdriver_geom
  implicit none
  integer, parameter :: nzone = 10, zsize = 2, nmem = 6, nobs = 8
contains
  ! i: 1-based state index (OAK passes startIndex(zi), rrsqrt.F90:368);
  ! element i sits at x = 0.5 i + 3, observation l at x = 1.25 l
  subroutine select_gauss(i, c, relevantObs)
    integer, intent(in) :: i
    real(8), intent(out) :: c(:)
    logical, intent(out) :: relevantObs(:)
    integer :: l
    real(8) :: d
    do l = 1, size(c)
      d = abs(0.5d0 * dble(i) + 3.0d0 - 1.25d0 * dble(l))
      relevantObs(l) = d <= 2.0d0
      c(l) = exp(-(d / 2.0d0)**2)
    end do
  end subroutine select_gauss
end module fdriver_geom

program oakg_fdriver
  use iso_c_binding
  use oakg_rrsqrt
  use fdriver_geom
  implicit none
  integer :: zoneSize(nzone), i, k, l, info, info2, info3, u, n
  real(c_double), allocatable, target :: xf(:), Hxf(:), yo(:), Sf(:,:), HSf(:,:), Rd(:), xa(:), Sa(:,:)
  real(c_double), allocatable, target :: xa2(:), Sa2(:,:), xa3(:), Sa3(:,:)
  real(c_double), allocatable, target :: zpos(:,:), cl(:), ml(:), ox(:), oy(:)
  character(len=1024) :: fname
  logical, allocatable :: excl(:)

  call get_command_argument(1, fname)
  n = nzone * zsize
  zoneSize = zsize
  allocate(xf(n), Hxf(nobs), yo(nobs), Sf(n, nmem), HSf(nobs, nmem), Rd(nobs), xa(n), Sa(n, nmem))
  allocate(xa2(n), Sa2(n, nmem), xa3(n), Sa3(n, nmem))
  allocate(zpos(4, nzone), cl(nzone), ml(nzone), ox(nobs), oy(nobs))
  do i = 1, n
    xf(i) = sin(0.3d0 * i)
    do k = 1, nmem
      Sf(i, k) = cos(0.7d0 * i * k) / nmem
    end do
  end do
  do l = 1, nobs
    yo(l) = 0.1d0 * l
    Hxf(l) = 0.005d0 * l * l
    Rd(l) = 0.1d0 + 0.01d0 * l
    do k = 1, nmem
      HSf(l, k) = cos(0.5d0 * l + 0.9d0 * k) / nmem
    end do
  end do
  xa = 0
  Sa = 0
  call locAnalysis_gpu(zoneSize, select_gauss, xf, Hxf, yo, Sf, HSf, Rd, xa, Sa, info=info)
  print '(a,i0)', 'locAnalysis_gpu info = ', info
  xa2 = 0
  Sa2 = 0
  info2 = info
  if (info == 0) &
    call locAnalysis_gpu(zoneSize, select_gauss, xf, Hxf, yo, Sf, HSf, Rd, xa2, Sa2, &
         localise_obs=.false., info=info2)
  print '(a,i0)', 'locAnalysis_gpu (localise_obs = .false.) info = ', info2
  ! geometry mode: the same selection (first element of zone i at
  ! x = 0.5 ((i-1) zsize + 1) + 3, observation l at x = 1.25 l, Cartesian,
  ! L = 2, cutoff 2) evaluated on the GPU
  zpos = 0
  do i = 1, nzone
    zpos(1, i) = 0.5d0 * dble((i - 1) * zsize + 1) + 3.0d0
  end do
  cl = 2.0d0
  ml = 2.0d0
  do l = 1, nobs
    ox(l) = 1.25d0 * dble(l)
  end do
  oy = 0
  xa3 = 0
  Sa3 = 0
  info3 = info2
  if (info2 == 0) &
    call locAnalysis_geo_gpu(zoneSize, zpos, cl, ml, 0, 1, ox, oy, oy, oy, xf, Hxf, yo, Sf, HSf, &
         Rd, xa3, Sa3, info=info3)
  print '(a,i0)', 'locAnalysis_geo_gpu info = ', info3
  ! the forecast report of Assim (assimilation.F90:3389), the last
  ! observation excluded
  allocate(excl(nobs))
  excl = .false.
  excl(nobs) = .true.
  if (info == 0) &
    call report_gpu(6, 'forecast.', 55000.5d0, count(.not. excl), Rd, HSf, yo - Hxf, excl)
  if (len_trim(fname) > 0) then
    open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
    write(u) int(info, 8), int(nzone, 8), int(zsize, 8), int(nmem, 8), int(nobs, 8), int(info2, 8), &
         int(info3, 8)
    write(u) xf, Hxf, yo, Sf, HSf, Rd, xa, Sa, xa2, Sa2, xa3, Sa3
    close(u)
  end if
  if (info /= 0) call exit(info)
  if (info2 /= 0) call exit(info2)
  if (info3 /= 0) call exit(info3)
end program oakg_fdriver
