! Fortran bindings of the C ABI in include/oakg.h (iso_c_binding).
!
! The types mirror the C structs field by field; array fields are C
! pointers (c_loc of contiguous Fortran arrays, or device addresses with
! OAKG_DEVICE_PTRS).  Indices inside the arrays are 0-based, as in C.
module oakg
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: OAKG_OK = 0, OAKG_EARG = 1, OAKG_ENAN = 2, OAKG_EHIP = 3, &
       OAKG_ENOMEM = 4, OAKG_EUNSUP = 5, OAKG_ECONV = 6
  integer(c_int), parameter :: OAKG_FP64 = 0, OAKG_MIXED = 1
  integer(c_int32_t), parameter :: OAKG_DEVICE_PTRS = 1
  integer(c_int), parameter :: OAKG_EIG_JACOBI = 0, OAKG_EIG_QL = 1
  ! launch shapes of the zone loop (oakg_set_tuning)
  integer(c_int), parameter :: OAKG_TUNE_SLOTS = 1, OAKG_TUNE_QL_ZONES = 2, OAKG_TUNE_GRAPH = 3, &
       OAKG_TUNE_ORDER = 6, OAKG_TUNE_TRIDIAG_SPLIT = 7, OAKG_TUNE_DUAL = 8, OAKG_TUNE_PERSIST = 9

  type, bind(C) :: oakg_zones
    integer(c_int64_t) :: nzones
    type(c_ptr) :: zone_ptr = c_null_ptr   ! int64 [nzones+1]
    type(c_ptr) :: zpos = c_null_ptr       ! double [nzones*4]
    type(c_ptr) :: corr_len = c_null_ptr   ! double [nzones]
    type(c_ptr) :: max_len = c_null_ptr    ! double [nzones]
    integer(c_int32_t) :: metric = 0
    integer(c_int32_t) :: loctype = 1
  end type oakg_zones

  type, bind(C) :: oakg_obs
    integer(c_int64_t) :: m
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr, t = c_null_ptr
    type(c_ptr) :: yo = c_null_ptr, rmse2 = c_null_ptr, excluded = c_null_ptr
  end type oakg_obs

  type, bind(C) :: oakg_csr
    integer(c_int64_t) :: m, n, nnz
    type(c_ptr) :: row_ptr = c_null_ptr, col = c_null_ptr, val = c_null_ptr
  end type oakg_csr

  interface
    integer(c_int) function oakg_version() bind(C, name="oakg_version")
      import :: c_int
    end function

    integer(c_int) function oakg_create(ctx, device, precision) bind(C, name="oakg_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device, precision
    end function

    integer(c_int) function oakg_destroy(ctx) bind(C, name="oakg_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function oakg_last_error(ctx) bind(C, name="oakg_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function oakg_set_eigensolver(ctx, method) bind(C, name="oakg_set_eigensolver")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: method
    end function

    integer(c_int) function oakg_set_tuning(ctx, key, value) bind(C, name="oakg_set_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: key, value
    end function

    integer(c_int) function oakg_local_analysis(ctx, zones, obs, N, xf, Hxf, Sf, ldSf, HSf, ldHSf, &
         xa, Sa, ldSa, amplitudes, localise_obs, flags) bind(C, name="oakg_local_analysis")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, oakg_zones, oakg_obs
      type(c_ptr), value :: ctx
      type(oakg_zones), intent(in) :: zones
      type(oakg_obs), intent(in) :: obs
      integer(c_int32_t), value :: N, localise_obs
      type(c_ptr), value :: xf, Hxf, Sf, HSf, xa, Sa, amplitudes
      integer(c_int64_t), value :: ldSf, ldHSf, ldSa
      integer(c_int32_t), value :: flags
    end function

    integer(c_int) function oakg_local_analysis_csr(ctx, nzones, zone_ptr, sel_ptr, sel_idx, sel_w, &
         obs, N, xf, Hxf, Sf, ldSf, HSf, ldHSf, xa, Sa, ldSa, amplitudes, localise_obs, flags) &
         bind(C, name="oakg_local_analysis_csr")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, oakg_obs
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nzones
      type(c_ptr), value :: zone_ptr, sel_ptr, sel_idx, sel_w
      type(oakg_obs), intent(in) :: obs
      integer(c_int32_t), value :: N, localise_obs
      type(c_ptr), value :: xf, Hxf, Sf, HSf, xa, Sa, amplitudes
      integer(c_int64_t), value :: ldSf, ldHSf, ldSa
      integer(c_int32_t), value :: flags
    end function

    integer(c_int) function oakg_ensemble_analysis(ctx, zones, obs, H, Hshift, N, Ef, ldEf, Ea, ldEa, &
         inflation, max_correction, max_correction_elem, flags) bind(C, name="oakg_ensemble_analysis")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double, oakg_zones, oakg_obs, oakg_csr
      type(c_ptr), value :: ctx
      type(oakg_zones), intent(in) :: zones
      type(oakg_obs), intent(in) :: obs
      type(oakg_csr), intent(in) :: H
      type(c_ptr), value :: Hshift, Ef, Ea, max_correction_elem
      integer(c_int32_t), value :: N
      integer(c_int64_t), value :: ldEf, ldEa
      real(c_double), value :: inflation, max_correction
      integer(c_int32_t), value :: flags
    end function

    integer(c_int) function oakg_report(ctx, m, N, yo_Hx, HS, ldHS, rmse2, excluded, nranges, &
         range_ptr, stats, flags) bind(C, name="oakg_report")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: m, ldHS
      integer(c_int32_t), value :: N, nranges
      type(c_ptr), value :: yo_Hx, HS, rmse2, excluded, range_ptr, stats
      integer(c_int32_t), value :: flags
    end function

    integer(c_int) function oakg_ensemble_report(ctx, obs, H, Hshift, N, E, ldE, nranges, &
         range_ptr, stats, flags) bind(C, name="oakg_ensemble_report")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, oakg_obs, oakg_csr
      type(c_ptr), value :: ctx
      type(oakg_obs), intent(in) :: obs
      type(oakg_csr), intent(in) :: H
      type(c_ptr), value :: Hshift, E, range_ptr, stats
      integer(c_int32_t), value :: N, nranges
      integer(c_int64_t), value :: ldE
      integer(c_int32_t), value :: flags
    end function
  end interface

contains

  ! Message of the last error of a context, as a Fortran string.
  function oakg_error_string(ctx) result(s)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: p(:)
    integer :: k
    type(c_ptr) :: cp
    cp = oakg_last_error(ctx)
    if (.not. c_associated(cp)) then
      s = ''
      return
    end if
    call c_f_pointer(cp, p, [4096])
    k = 0
    do while (k < 4096)
      if (p(k + 1) == c_null_char) exit
      k = k + 1
    end do
    allocate(character(len=k) :: s)
    s = transfer(p(1:k), s)
  end function oakg_error_string
end module oakg
