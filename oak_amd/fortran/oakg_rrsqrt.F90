! Drop-in for OAK's locAnalysis (rrsqrt.F90:433-457) on the GPU.
!
! locAnalysis_gpu keeps locAnalysis' argument list; R is passed as its
! diagonal (the DiagCovar component R%D, covariance.F90:70-79), since OAK's
! local scheme uses a diagonal R.  The caller's selectObservations callback is
! evaluated per zone on the host (zone i = 1..size(zoneSize), as
! locAnalysisIncrement does, rrsqrt.F90:358-418), which keeps any selection
! exact; the relevant observations and their weights go to the GPU as CSR.
! Like locAnalysis, it may be called by every thread of an OpenMP parallel
! region: the work runs in the master thread and ends with a barrier.
module oakg_rrsqrt
  use iso_c_binding
  use oakg
  implicit none
  private
  public :: locAnalysis_gpu, oakg_rrsqrt_context

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
    integer :: nz, m, n, iz, k, p, nn
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
    ! selection: count, then fill (0-based observation indices)
    sel_ptr(1) = 0
    do iz = 1, nz
      call selectObservations(iz, wz, rel)
      sel_ptr(iz + 1) = sel_ptr(iz) + count(rel)
    end do
    nn = int(sel_ptr(nz + 1))
    allocate(sel_idx(max(nn, 1)), sel_w(max(nn, 1)))
    p = 0
    do iz = 1, nz
      call selectObservations(iz, wz, rel)
      do k = 1, m
        if (rel(k)) then
          p = p + 1
          sel_idx(p) = int(k - 1, c_int32_t)
          sel_w(p) = wz(k)
        end if
      end do
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
  end subroutine locAnalysis_gpu
end module oakg_rrsqrt
