!> The MI355X plug-in of the reference's static structure-factor driver.
!!
!! A maintainer adds this file, with the binding module sqc_hip.f90, to the
!! reference's src/ and selects it in app/main.f90:168-184 in place of
!! cufinufft_structure_factor_t.  It extends the reference's abstract
!! structure_factor_t (src/sqc_structure_factor.f90:71-192) and has the shape
!! of the CUDA plug-in (src/sqc_gpu.f90:42-71):
!!   method_setup      plan from the first frame (gpu_setup, src/sqc_gpu.f90:108-210)
!!   accumulate_frame  the frame's Cartesian positions to the device
!!                     (gpu_accumulate, src/sqc_gpu.f90:213-362)
!!   prepare_output    the device accumulators into num, partial_num and gnum,
!!                     which write_results reads (src/sqc_structure_factor.f90:647-654,
!!                     app/main.f90:286)
!! The base class's configure (sf_configure) has already built den,
!! shell_count and the mode tables by the same rules as the library.  Weights
!! come from the reference's own tables: b through scheme%amplitude, and the
!! IT92 coefficients of each species from species_xray_a/b/c of
!! src/sqc_element_data.f90 at the species index the mapping resolved.
module sqc_hip_sf
   use, intrinsic :: iso_c_binding
   use sqc_kinds
   use sqc_structure_factor, only: structure_factor_t, sf_alloc_partials
   use sqc_dump, only: frame_t
   use sqc_weights, only: weight_scheme_t, weight_unit, weight_neutron, weight_xray
   use sqc_element_data, only: n_species, species_xray_a, species_xray_b, species_xray_c
   use sqc_hip
   implicit none
   private
   public :: hip_structure_factor_t

   type, extends(structure_factor_t) :: hip_structure_factor_t
      type(c_ptr) :: plan = c_null_ptr
      !> --gpu-id and --precision single of the driver's options
      integer :: device = 0
      logical :: single_precision = .false.
      integer :: nspecies = 0
      integer(ik), allocatable :: species_type(:)
   contains
      procedure :: method_setup => hip_setup
      procedure :: accumulate_frame => hip_accumulate
      procedure :: prepare_output => hip_prepare_output
      final :: hip_finalize
   end type hip_structure_factor_t

contains

   subroutine hip_setup(self, frame, scheme, ierr, message)
      class(hip_structure_factor_t), intent(inout) :: self
      type(frame_t), intent(in) :: frame
      type(weight_scheme_t), intent(in) :: scheme
      integer, intent(out) :: ierr
      character(len=*), intent(out) :: message
      type(sqc_hip_config_t) :: cfg
      integer(c_int32_t), allocatable, target :: species_of(:)
      real(c_double), allocatable, target :: amp(:), coef(:, :)
      integer(ik) :: t
      integer :: a, j, sidx
      integer(c_int) :: rc
      logical, allocatable :: present(:)

      ierr = 0
      message = ''
      if (.not. sqc_hip_layout_ok()) then            ! src/sqc_cufinufft.f90:277-290
         ierr = 1
         message = 'libsqc_hip uses an incompatible sqc_hip_config layout'
         return
      end if
      if (self%xrd_enabled) then
         ierr = 1
         message = '--xrd on the HIP path: use sqcalc_hip static --xrd'
         return
      end if
      ! species = type ids present, ascending (src/sqc_structure_factor.f90:1127-1144)
      allocate (present(maxval(frame%type_id)))
      present = .false.
      present(frame%type_id) = .true.
      self%species_type = pack([(t, t = 1_ik, size(present, kind=ik))], present)
      self%nspecies = size(self%species_type)
      allocate (species_of(frame%natoms), amp(self%nspecies), coef(9, self%nspecies))
      do j = 1, int(frame%natoms)
         species_of(j) = int(findloc(self%species_type, frame%type_id(j), 1) - 1, c_int32_t)
      end do
      amp = 1.0_rk
      coef = 0.0_rk
      do a = 1, self%nspecies
         t = self%species_type(a)
         select case (scheme%kind)
         case (weight_neutron)
            amp(a) = scheme%amplitude(t, 0.0_rk)
         case (weight_xray)
            sidx = 0
            if (t <= size(scheme%species_index, kind=ik)) sidx = scheme%species_index(t)
            if (sidx < 1 .or. sidx > n_species) then
               ierr = 1
               message = 'no X-ray form factor for '//trim(scheme%species(t))
               return
            end if
            coef(1:4, a) = species_xray_a(:, sidx)
            coef(5:8, a) = species_xray_b(:, sidx)
            coef(9, a) = species_xray_c(sidx)
         end select
      end do
      ! the pair columns and their labels, as the CUDA plug-in allocates them
      if (self%partials) call sf_alloc_partials(self, scheme)
      cfg%device = int(self%device, c_int32_t)
      cfg%precision = merge(sqc_hip_f32, sqc_hip_f64, self%single_precision)
      select case (scheme%kind)
      case (weight_neutron)
         cfg%weight_kind = sqc_hip_weight_neutron
      case (weight_xray)
         cfg%weight_kind = sqc_hip_weight_xray
      case default
         cfg%weight_kind = sqc_hip_weight_unit
      end select
      cfg%norm = int(self%norm, c_int32_t)   ! norm_mean/self/natom = 1/2/3 (:59-61)
      cfg%partials = merge(1_c_int32_t, 0_c_int32_t, self%partials)
      cfg%want_grid = merge(1_c_int32_t, 0_c_int32_t, self%want_grid)
      cfg%nq = int(self%nq, c_int32_t)
      cfg%qmin = self%qmin
      cfg%qmax = self%qmax
      cfg%eps = self%eps
      cfg%cell = reshape(frame%cell%a, [9])
      cfg%origin = frame%cell%origin
      cfg%natoms = int(frame%natoms, c_int64_t)
      cfg%nspecies = int(self%nspecies, c_int32_t)
      cfg%species_of = c_loc(species_of)
      cfg%amp_const = c_loc(amp)
      cfg%xray_coef = c_loc(coef)
      rc = sqc_hip_plan_create(self%plan, cfg)     ! replaces cufinufft_makeplan
      if (rc /= sqc_hip_ok .and. rc /= sqc_hip_warn_eps_too_small) then
         ierr = 1
         message = sqc_hip_message(rc)//': '//sqc_hip_detail()
      end if
   end subroutine hip_setup

   subroutine hip_accumulate(self, frame, scheme, ierr, message)
      class(hip_structure_factor_t), intent(inout) :: self
      type(frame_t), intent(in) :: frame
      type(weight_scheme_t), intent(in) :: scheme
      integer, intent(out) :: ierr
      character(len=*), intent(out) :: message
      integer(c_int) :: rc
      ierr = 0
      message = ''
      ! wrap, spread, FFT, deconvolution, species combination and shell sums
      ! all run on the device; nothing comes back per frame
      rc = sqc_hip_accumulate(self%plan, frame%pos, 1_c_int64_t)
      if (rc /= sqc_hip_ok) then
         ierr = 1
         message = sqc_hip_message(rc)//': '//sqc_hip_detail()
         return
      end if
      self%nframes = self%nframes + 1
   end subroutine hip_accumulate

   subroutine hip_prepare_output(self, scheme, ierr, message)
      class(hip_structure_factor_t), intent(inout) :: self
      type(weight_scheme_t), intent(in) :: scheme
      integer, intent(out) :: ierr
      character(len=*), intent(out) :: message
      real(c_double), allocatable, target :: part(:, :, :), gk(:)
      integer(c_int64_t) :: nfr
      integer(c_int) :: rc
      integer :: a, b
      type(c_ptr) :: gptr
      ierr = 0
      message = ''
      allocate (part(self%nq, self%nspecies, self%nspecies), gk(max(self%nmodes, 1_lk)))
      gptr = c_null_ptr
      if (self%want_grid) gptr = c_loc(gk)
      rc = sqc_hip_fetch(self%plan, self%num, c_loc(part), gptr, nfr)
      if (rc /= sqc_hip_ok) then
         ierr = 1
         message = sqc_hip_message(rc)//': '//sqc_hip_detail()
         return
      end if
      self%nframes = int(nfr, lk)
      if (self%partials) then
         self%partial_num = 0.0_rk
         do a = 1, self%nspecies                  ! species index -> LAMMPS type id
            do b = a, self%nspecies
               self%partial_num(self%species_type(a), self%species_type(b), :) = part(:, b, a)
            end do
         end do
      end if
      if (self%want_grid) self%gnum(self%gidx) = gk(1:self%nmodes)
   end subroutine hip_prepare_output

   subroutine hip_finalize(self)
      type(hip_structure_factor_t), intent(inout) :: self
      integer(c_int) :: rc
      if (c_associated(self%plan)) rc = sqc_hip_destroy(self%plan)
      self%plan = c_null_ptr
   end subroutine hip_finalize

end module sqc_hip_sf
