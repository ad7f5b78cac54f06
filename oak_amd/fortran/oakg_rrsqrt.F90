! Drop-in for OAK's locAnalysis (rrsqrt.F90:433-457) on the GPU.
!
! locAnalysis_gpu keeps locAnalysis' argument list; R is passed as its
! diagonal (the DiagCovar component R%D, covariance.F90:70-79), since OAK's
! local scheme uses a diagonal R.  The caller's selectObservations callback is
! evaluated once per zone on the host with the 1-based state index of the
! zone's first element, as locAnalysisIncrement calls it
! (startIndex(zi), rrsqrt.F90:358-368), which keeps any selection exact; the
! relevant observations and their weights go to the GPU as CSR.  With
! localise_obs = .false., a zone with at least one relevant observation uses
! all observations with their weights (rrsqrt.F90:373-385).
! Like locAnalysis, it may be called by every thread of an OpenMP parallel
! region: the work runs in the master thread and ends with a barrier.
module oakg_rrsqrt
  use iso_c_binding
  use oakg
  implicit none
  private
  public :: locAnalysis_gpu, locAnalysis_geo_gpu, report_gpu, oakg_rrsqrt_context

  type(c_ptr), save :: ctx = c_null_ptr

contains

  ! The context (device 0, FP64), created on first use.
  function oakg_rrsqrt_context() result(c)
    type(c_ptr) :: c
    integer(c_int) :: rc
    if (.not. c_associated(ctx)) then
      rc = oakg_create(ctx, 0_c_int, OAKG_FP64)
      if (rc /= OAKG_OK) then
        ctx = c_null_ptr
      end if
    end if
    c = ctx
  end function oakg_rrsqrt_context

  subroutine locAnalysis_gpu(zoneSize, selectObservations, xf, Hxf, yo, Sf, HSf, Rdiag, xa, Sa, &
       amplitudes, localise_obs, info)
    interface
      subroutine selectObservations(i, c, relevantObs)
        integer, intent(in) :: i
        real(8), intent(out) :: c(:)
        logical, intent(out) :: relevantObs(:)
      end subroutine selectObservations
    end interface
    integer, intent(in) :: zoneSize(:)
    real(c_double), intent(in), target, contiguous :: xf(:), Hxf(:), yo(:), Sf(:,:), HSf(:,:), Rdiag(:)
    real(c_double), intent(out), target, contiguous :: xa(:), Sa(:,:)
    real(c_double), intent(out), target, optional, contiguous :: amplitudes(:,:)
    logical, intent(in), optional :: localise_obs
    integer, intent(out), optional :: info

    integer(c_int64_t), allocatable, target :: zone_ptr(:), sel_ptr(:)
    integer(c_int32_t), allocatable, target :: sel_idx(:)
    real(c_double), allocatable, target :: sel_w(:), wz(:)
    logical, allocatable :: rel(:)
    type(oakg_obs) :: obs
    type(c_ptr) :: c, pampl
    integer(c_int) :: rc
    integer :: nz, m, n, iz, k, l
    integer(c_int64_t) :: p, nn
    integer(c_int32_t) :: loc

!$omp master
    nz = size(zoneSize)
    m = size(yo)
    n = size(xf)
    loc = 1
    if (present(localise_obs)) then
      if (.not. localise_obs) loc = 0
    end if
    allocate(zone_ptr(nz + 1), sel_ptr(nz + 1), wz(m), rel(m))
    zone_ptr(1) = 0
    do iz = 1, nz
      zone_ptr(iz + 1) = zone_ptr(iz) + zoneSize(iz)
    end do
    ! selection: one callback per zone with the zone's first state index,
    ! its relevant observations appended to the CSR lists (0-based
    ! observation indices; the lists grow by doubling)
    nn = max(16, m)
    allocate(sel_idx(nn), sel_w(nn))
    sel_ptr(1) = 0
    p = 0
    do iz = 1, nz
      call selectObservations(int(zone_ptr(iz)) + 1, wz, rel)
      k = count(rel)
      ! localise_obs = .false.: all observations (with their weights) as soon
      ! as one is relevant (rrsqrt.F90:373-385)
      if (loc == 0 .and. k > 0) then
        rel = .true.
        k = m
      end if
      if (p + k > nn) call grow(max(2 * nn, p + k))
      do l = 1, m
        if (rel(l)) then
          p = p + 1
          sel_idx(p) = int(l - 1, c_int32_t)
          sel_w(p) = wz(l)
        end if
      end do
      sel_ptr(iz + 1) = p
    end do
    obs%m = m
    obs%yo = c_loc(yo)
    obs%rmse2 = c_loc(Rdiag)
    pampl = c_null_ptr
    if (present(amplitudes)) pampl = c_loc(amplitudes)
    c = oakg_rrsqrt_context()
    if (.not. c_associated(c)) then
      rc = OAKG_EHIP
    else
      rc = oakg_local_analysis_csr(c, int(nz, c_int64_t), c_loc(zone_ptr), c_loc(sel_ptr), &
           c_loc(sel_idx), c_loc(sel_w), obs, int(size(Sf, 2), c_int32_t), c_loc(xf), c_loc(Hxf), &
           c_loc(Sf), int(n, c_int64_t), c_loc(HSf), int(m, c_int64_t), c_loc(xa), c_loc(Sa), &
           int(n, c_int64_t), pampl, loc, 0_c_int32_t)
    end if
    if (present(info)) then
      info = rc
    else if (rc /= OAKG_OK) then
      if (c_associated(c)) write(0, *) 'locAnalysis_gpu: ', oakg_error_string(c)
      ! OAK stops on a NaN amplitude (rrsqrt.F90:145-149); every error stops here
      error stop 1
    end if
!$omp end master
!$omp barrier
  contains
    subroutine grow(cap)
      integer(c_int64_t), intent(in) :: cap
      integer(c_int32_t), allocatable, target :: i2(:)
      real(c_double), allocatable, target :: w2(:)
      allocate(i2(cap), w2(cap))
      i2(1:p) = sel_idx(1:p)
      w2(1:p) = sel_w(1:p)
      call move_alloc(i2, sel_idx)
      call move_alloc(w2, sel_w)
      nn = cap
    end subroutine grow
  end subroutine locAnalysis_gpu

  ! locAnalysis with OAK's own selectObservations (assimilation.F90:3683-3771)
  ! evaluated on the GPU for every zone at once, instead of a callback per
  ! zone on the host: the drop-in for the call in Assim
  ! (assimilation.F90:3235-3236).  The assimilation module state that OAK's
  ! selectObservations reads is passed flattened:
  !   zpos(4, nzones)  x, y, z, t of each zone's first element, i.e. of
  !                    zoneIndex(startIndex(zi)) through ind2submv and getCoord
  !                    (:3713-3740)
  !   corrLen, maxLen  hCorrLengthToObs / hMaxCorrLengthToObs of that element
  !                    (:414-415, 3741-3742), one per zone
  !   metric, loctype  metrictype (0 Cartesian, 1 spherical, 2 approximate)
  !                    and loctype (1 horizontal, 2 |dz|, 3 |dt|)
  !   obsX, obsY, obsZ, obsT  obsGridX/Y/Z/T (:228, 3159-3167)
  ! The other arguments are those of locAnalysis_gpu.
  subroutine locAnalysis_geo_gpu(zoneSize, zpos, corrLen, maxLen, metric, loctype, obsX, obsY, &
       obsZ, obsT, xf, Hxf, yo, Sf, HSf, Rdiag, xa, Sa, amplitudes, localise_obs, info)
    integer, intent(in) :: zoneSize(:), metric, loctype
    real(c_double), intent(in), target, contiguous :: zpos(:,:), corrLen(:), maxLen(:)
    real(c_double), intent(in), target, contiguous :: obsX(:), obsY(:), obsZ(:), obsT(:)
    real(c_double), intent(in), target, contiguous :: xf(:), Hxf(:), yo(:), Sf(:,:), HSf(:,:), Rdiag(:)
    real(c_double), intent(out), target, contiguous :: xa(:), Sa(:,:)
    real(c_double), intent(out), target, optional, contiguous :: amplitudes(:,:)
    logical, intent(in), optional :: localise_obs
    integer, intent(out), optional :: info

    integer(c_int64_t), allocatable, target :: zone_ptr(:)
    type(oakg_zones) :: zones
    type(oakg_obs) :: obs
    type(c_ptr) :: c, pampl
    integer(c_int) :: rc
    integer :: nz, m, n, iz
    integer(c_int32_t) :: loc

!$omp master
    nz = size(zoneSize)
    m = size(yo)
    n = size(xf)
    loc = 1
    if (present(localise_obs)) then
      if (.not. localise_obs) loc = 0
    end if
    allocate(zone_ptr(nz + 1))
    zone_ptr(1) = 0
    do iz = 1, nz
      zone_ptr(iz + 1) = zone_ptr(iz) + zoneSize(iz)
    end do
    zones%nzones = nz
    zones%zone_ptr = c_loc(zone_ptr)
    zones%zpos = c_loc(zpos)
    zones%corr_len = c_loc(corrLen)
    zones%max_len = c_loc(maxLen)
    zones%metric = int(metric, c_int32_t)
    zones%loctype = int(loctype, c_int32_t)
    obs%m = m
    obs%x = c_loc(obsX)
    obs%y = c_loc(obsY)
    obs%z = c_loc(obsZ)
    obs%t = c_loc(obsT)
    obs%yo = c_loc(yo)
    obs%rmse2 = c_loc(Rdiag)
    pampl = c_null_ptr
    if (present(amplitudes)) pampl = c_loc(amplitudes)
    c = oakg_rrsqrt_context()
    if (.not. c_associated(c)) then
      rc = OAKG_EHIP
    else
      rc = oakg_local_analysis(c, zones, obs, int(size(Sf, 2), c_int32_t), c_loc(xf), c_loc(Hxf), &
           c_loc(Sf), int(n, c_int64_t), c_loc(HSf), int(m, c_int64_t), c_loc(xa), c_loc(Sa), &
           int(n, c_int64_t), pampl, loc, 0_c_int32_t)
    end if
    if (present(info)) then
      info = rc
    else if (rc /= OAKG_OK) then
      if (c_associated(c)) write(0, *) 'locAnalysis_geo_gpu: ', oakg_error_string(c)
      error stop 1
    end if
!$omp end master
!$omp barrier
  end subroutine locAnalysis_geo_gpu

  ! report (assimilation.F90:3584-3628) with the statistics computed on the
  ! GPU: the same arguments, with R given by its diagonal Rdiag (R%D), and the
  ! same log lines.  ingrid is the divisor, as in OAK (the caller passes
  ! count(.not.exclude_obs)); the excluded observations drop out of the sums
  ! as with OAK's DCDCovar-wrapped R (assimilation.F90:3086-3092).
  subroutine report_gpu(unit, prefix, mjd, ingrid, Rdiag, HS, yo_Hx, exclude_obs)
    integer, intent(in) :: unit
    character(len=*), intent(in) :: prefix
    real(8), intent(in) :: mjd
    integer, intent(in) :: ingrid
    real(c_double), intent(in), target, contiguous :: Rdiag(:), HS(:,:), yo_Hx(:)
    logical, intent(in) :: exclude_obs(:)
    character(len=*), parameter :: form = '(A,A,E15.7)'
    integer(c_int8_t), allocatable, target :: ex(:)
    real(c_double), target :: st(5)
    real(c_double) :: g
    type(c_ptr) :: c
    integer(c_int) :: rc
    allocate(ex(size(yo_Hx)))
    ex = 0
    where (exclude_obs) ex = 1
    c = oakg_rrsqrt_context()
    rc = OAKG_EHIP
    if (c_associated(c)) &
      rc = oakg_report(c, int(size(yo_Hx), c_int64_t), int(size(HS, 2), c_int32_t), &
           c_loc(yo_Hx), c_loc(HS), int(max(size(HS, 1), 1), c_int64_t), c_loc(Rdiag), c_loc(ex), &
           1_c_int32_t, c_null_ptr, c_loc(st), 0_c_int32_t)
    if (rc /= OAKG_OK) then
      if (c_associated(c)) write(0, *) 'report_gpu: ', oakg_error_string(c)
      error stop 1
    end if
    ! the library divides by count(.not.exclude_obs); another ingrid of the
    ! caller rescales the sums
    g = real(ingrid, c_double)
    if (g /= st(1)) then
      st(2) = sqrt(st(2)**2 * st(1) / g)
      st(3) = st(3) * st(1) / g
      st(4) = sqrt(st(4)**2 * st(1) / g)
      st(5) = sqrt(st(5)**2 * st(1) / g)
    end if
    write(unit,'(A,A,F15.4)') prefix,'mjd                      ',mjd
    write(unit,form) prefix,'rms_yo-Hx                ',st(2)
    write(unit,form) prefix,'bias_yo-Hx               ',st(3)
    write(unit,form) prefix,'rms_invsqrtR_yo-Hx       ',st(4)
    write(unit,form) prefix,'projection_yo-Hx_into_HSf',st(5)
  end subroutine report_gpu
end module oakg_rrsqrt
