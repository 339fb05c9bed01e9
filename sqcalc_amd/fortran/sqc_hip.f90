!> Fortran 2008 binding of libsqc_hip (include/sqc_hip.h) over iso_c_binding.
!!
!! Replaces the reference's cuFINUFFT binding module (src/sqc_cufinufft.f90:
!! interfaces at :64-140, CUDA runtime calls at :143-174, opts layout check at
!! :277-290): instead of a transform plan plus host-side binning, the library
!! takes positions and returns the shell accumulators.  The derived types
!! below are interoperable with the C structs and must keep their field order.
module sqc_hip
   use, intrinsic :: iso_c_binding
   implicit none
   private

   integer(c_int32_t), parameter, public :: sqc_hip_abi_version = 2

   integer(c_int), parameter, public :: sqc_hip_ok = 0
   integer(c_int), parameter, public :: sqc_hip_warn_eps_too_small = 1
   integer(c_int), parameter, public :: sqc_hip_err_maxnalloc = 2
   integer(c_int), parameter, public :: sqc_hip_err_alloc = 11
   integer(c_int), parameter, public :: sqc_hip_err_bad_arg = 20
   integer(c_int), parameter, public :: sqc_hip_err_abi = 30
   integer(c_int), parameter, public :: sqc_hip_err_no_device = 31
   integer(c_int), parameter, public :: sqc_hip_err_hip = 32
   integer(c_int), parameter, public :: sqc_hip_err_no_modes = 33

   integer(c_int32_t), parameter, public :: sqc_hip_f64 = 0, sqc_hip_f32 = 1
   integer(c_int32_t), parameter, public :: sqc_hip_weight_unit = 0, &
                                            sqc_hip_weight_neutron = 1, &
                                            sqc_hip_weight_xray = 2
   integer(c_int32_t), parameter, public :: sqc_hip_norm_mean = 1, sqc_hip_norm_self = 2, &
                                            sqc_hip_norm_natom = 3

   !> struct sqc_hip_config
   type, bind(C), public :: sqc_hip_config_t
      integer(c_int32_t) :: abi_version = sqc_hip_abi_version
      integer(c_int32_t) :: device = 0
      integer(c_int32_t) :: precision = sqc_hip_f64
      integer(c_int32_t) :: weight_kind = sqc_hip_weight_unit
      integer(c_int32_t) :: norm = sqc_hip_norm_mean
      integer(c_int32_t) :: partials = 1
      integer(c_int32_t) :: want_grid = 0
      integer(c_int32_t) :: nq = 500
      real(c_double) :: qmin = 0.0_c_double
      real(c_double) :: qmax = 20.0_c_double
      real(c_double) :: eps = 1.0e-9_c_double
      real(c_double) :: upsampfac = 0.0_c_double
      real(c_double) :: cell(9) = 0.0_c_double
      real(c_double) :: origin(3) = 0.0_c_double
      integer(c_int64_t) :: natoms = 0
      integer(c_int32_t) :: nspecies = 0
      integer(c_int32_t) :: mode_partials = 0
      type(c_ptr) :: species_of = c_null_ptr
      type(c_ptr) :: amp_const = c_null_ptr
      type(c_ptr) :: xray_coef = c_null_ptr
      integer(c_int32_t) :: tune = 0
      integer(c_int32_t) :: max_batch = 0
   end type sqc_hip_config_t

   !> struct sqc_hip_info
   type, bind(C), public :: sqc_hip_info_t
      integer(c_int64_t) :: modes(3)
      integer(c_int64_t) :: nmodes
      integer(c_int64_t) :: nhalf
      integer(c_int64_t) :: fine(3)
      integer(c_int32_t) :: nspread
      integer(c_int32_t) :: nfields
      real(c_double) :: beta
      real(c_double) :: upsampfac
      integer(c_int64_t) :: device_bytes
      integer(c_int64_t) :: nlines
      integer(c_int32_t) :: tile(3)
      integer(c_int32_t) :: ncomp
      integer(c_int64_t) :: rho_len
      integer(c_int32_t) :: shell_sums_in_hbm
      integer(c_int32_t) :: x_lines
   end type sqc_hip_info_t

   public :: sqc_hip_device_count, sqc_hip_plan_create, sqc_hip_plan_info, sqc_hip_set_tiles, &
             sqc_hip_shells, sqc_hip_modes, sqc_hip_accumulate, sqc_hip_fetch, sqc_hip_fetch_mode_partials, sqc_hip_reset, &
             sqc_hip_gradient, sqc_hip_destroy, sqc_hip_config_size
   public :: sqc_hip_message, sqc_hip_detail, sqc_hip_layout_ok

   interface
      integer(c_int) function sqc_hip_device_count(count) bind(C, name='sqc_hip_device_count')
         import :: c_int
         integer(c_int), intent(out) :: count
      end function sqc_hip_device_count

      integer(c_int) function sqc_hip_plan_create(plan, cfg) bind(C, name='sqc_hip_plan_create')
         import :: c_int, c_ptr, sqc_hip_config_t
         type(c_ptr), intent(out) :: plan
         type(sqc_hip_config_t), intent(in) :: cfg
      end function sqc_hip_plan_create

      integer(c_int) function sqc_hip_plan_info(plan, info) bind(C, name='sqc_hip_plan_info')
         import :: c_int, c_ptr, sqc_hip_info_t
         type(c_ptr), value :: plan
         type(sqc_hip_info_t), intent(out) :: info
      end function sqc_hip_plan_info

      ! pin the spread tile, y-pass column chunk and x-pass line group (the
      ! cuFINUFFT gpu_binsize options of src/sqc_cufinufft.f90:45-53)
      integer(c_int) function sqc_hip_set_tiles(plan, spread_x, spread_y, y_chunk, x_lines) &
         bind(C, name='sqc_hip_set_tiles')
         import :: c_int, c_ptr
         type(c_ptr), value :: plan
         integer(c_int), value :: spread_x, spread_y, y_chunk, x_lines
      end function sqc_hip_set_tiles

      integer(c_int) function sqc_hip_shells(plan, den, shell_count) &
         bind(C, name='sqc_hip_shells')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: plan
         real(c_double), intent(out) :: den(*)
         integer(c_int64_t), intent(out) :: shell_count(*)
      end function sqc_hip_shells

      integer(c_int) function sqc_hip_modes(plan, hkl, qvec, gidx) bind(C, name='sqc_hip_modes')
         import :: c_int, c_ptr
         type(c_ptr), value :: plan, hkl, qvec, gidx
      end function sqc_hip_modes

      integer(c_int) function sqc_hip_accumulate(plan, pos, nframes) &
         bind(C, name='sqc_hip_accumulate')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: plan
         real(c_double), intent(in) :: pos(*)
         integer(c_int64_t), value :: nframes
      end function sqc_hip_accumulate

      integer(c_int) function sqc_hip_fetch(plan, num, partial, gnum, nframes) &
         bind(C, name='sqc_hip_fetch')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: plan
         real(c_double), intent(out) :: num(*)
         type(c_ptr), value :: partial, gnum
         integer(c_int64_t), intent(out) :: nframes
      end function sqc_hip_fetch

      integer(c_int) function sqc_hip_fetch_mode_partials(plan, gpart) &
         bind(C, name='sqc_hip_fetch_mode_partials')
         import :: c_int, c_ptr
         type(c_ptr), value :: plan, gpart
      end function sqc_hip_fetch_mode_partials

      integer(c_int) function sqc_hip_reset(plan) bind(C, name='sqc_hip_reset')
         import :: c_int, c_ptr
         type(c_ptr), value :: plan
      end function sqc_hip_reset

      integer(c_int) function sqc_hip_gradient(plan, pos, target, wq, loss, grad) &
         bind(C, name='sqc_hip_gradient')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: plan
         real(c_double), intent(in) :: pos(*), target(*)
         type(c_ptr), value :: wq
         real(c_double), intent(out) :: loss
         real(c_double), intent(out) :: grad(*)
      end function sqc_hip_gradient

      integer(c_int) function sqc_hip_destroy(plan) bind(C, name='sqc_hip_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: plan
      end function sqc_hip_destroy

      integer(c_size_t) function sqc_hip_config_size() bind(C, name='sqc_hip_config_size')
         import :: c_size_t
      end function sqc_hip_config_size

      type(c_ptr) function c_error_string(code) bind(C, name='sqc_hip_error_string')
         import :: c_ptr, c_int
         integer(c_int), value :: code
      end function c_error_string

      type(c_ptr) function c_last_error() bind(C, name='sqc_hip_last_error')
         import :: c_ptr
      end function c_last_error

      integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function c_strlen
   end interface

contains

   !> Text of a status code (cufinufft_error_text of src/sqc_cufinufft.f90:243-274).
   function sqc_hip_message(code) result(text)
      integer(c_int), intent(in) :: code
      character(len=:), allocatable :: text
      text = c_to_f(c_error_string(code))
   end function sqc_hip_message

   !> Detail of the last failure on this thread.
   function sqc_hip_detail() result(text)
      character(len=:), allocatable :: text
      text = c_to_f(c_last_error())
   end function sqc_hip_detail

   !> Layout check of the config struct against the library, the counterpart
   !! of the reference's opts ABI check (src/sqc_cufinufft.f90:277-290).
   logical function sqc_hip_layout_ok()
      type(sqc_hip_config_t), target :: probe
      sqc_hip_layout_ok = c_sizeof(probe) == sqc_hip_config_size()
   end function sqc_hip_layout_ok

   function c_to_f(p) result(text)
      type(c_ptr), intent(in) :: p
      character(len=:), allocatable :: text
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      if (.not. c_associated(p)) then
         text = ''
         return
      end if
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate (character(len=n) :: text)
      do i = 1, n
         text(i:i) = chars(i)
      end do
   end function c_to_f

end module sqc_hip
