!> Species labels of -m TYPE:LABEL and their scattering data, through the
!! library's lookup (sqc_hip_species_lookup / sqc_hip_species_table in
!! include/sqc_hip.h).
!!
!! The library restates the reference's label rules (src/sqc_elements.f90:
!! 241-411: aliases, any case, ions "O2-" / "Na1+" / "Fe+", valence states
!! "Cval") over the data of src/sqc_element_data.f90 (104 elements with
!! their bound coherent neutron lengths, 213 X-ray species with IT92
!! coefficients).  An ion or valence label takes its element's neutron
!! length (src/sqc_weights.f90:136-170).
module sqc_hip_elements
   use, intrinsic :: iso_c_binding
   use, intrinsic :: iso_fortran_env, only: real64
   implicit none
   private
   public :: sqc_hip_species_info_t, resolve_species, species_table_size, species_table_row, &
             info_label

   integer, parameter :: rk = real64

   !> struct sqc_hip_species_info
   type, bind(C) :: sqc_hip_species_info_t
      character(kind=c_char) :: label(16)
      character(kind=c_char) :: symbol(4)
      integer(c_int32_t) :: z
      integer(c_int32_t) :: charge
      integer(c_int32_t) :: valence
      integer(c_int32_t) :: has_neutron
      integer(c_int32_t) :: has_xray
      integer(c_int32_t) :: reserved
      real(c_double) :: neutron_b
      real(c_double) :: xray_coef(9)
   end type sqc_hip_species_info_t

   interface
      integer(c_int) function c_species_lookup(text, info) bind(C, name='sqc_hip_species_lookup')
         import :: c_int, c_char, sqc_hip_species_info_t
         character(kind=c_char), intent(in) :: text(*)
         type(sqc_hip_species_info_t), intent(out) :: info
      end function c_species_lookup

      integer(c_int) function c_species_table(kind, index, info, count) &
         bind(C, name='sqc_hip_species_table')
         import :: c_int, c_int32_t, c_ptr
         integer(c_int32_t), value :: kind, index
         type(c_ptr), value :: info, count
      end function c_species_table

      type(c_ptr) function c_species_error() bind(C, name='sqc_hip_species_error')
         import :: c_ptr
      end function c_species_error

      integer(c_size_t) function c_strlen(p) bind(C, name='strlen')
         import :: c_size_t, c_ptr
         type(c_ptr), value :: p
      end function c_strlen
   end interface

contains

   !> Resolve a label (table_find_species): ierr 0, or the library's code
   !! (34 unknown element, 35 no such charge state) with its message.
   subroutine resolve_species(text, info, ierr, message)
      character(len=*), intent(in) :: text
      type(sqc_hip_species_info_t), intent(out) :: info
      integer, intent(out) :: ierr
      character(len=:), allocatable, intent(out) :: message
      ierr = int(c_species_lookup(trim(text)//c_null_char, info))
      message = ''
      if (ierr /= 0) message = c_text(c_species_error())
   end subroutine resolve_species

   !> Rows of the element table (kind 0) or of the X-ray species table (kind 1).
   integer function species_table_size(kind) result(n)
      integer, intent(in) :: kind
      integer(c_int32_t), target :: cnt
      integer(c_int) :: rc
      rc = c_species_table(int(kind, c_int32_t), 0_c_int32_t, c_null_ptr, c_loc(cnt))
      n = int(cnt)
      if (rc /= 0) n = 0
   end function species_table_size

   subroutine species_table_row(kind, index, info)
      integer, intent(in) :: kind, index
      type(sqc_hip_species_info_t), target, intent(out) :: info
      integer(c_int) :: rc
      rc = c_species_table(int(kind, c_int32_t), int(index, c_int32_t), c_loc(info), c_null_ptr)
   end subroutine species_table_row

   !> The canonical label of a resolved species, as Fortran text.
   function info_label(info) result(text)
      type(sqc_hip_species_info_t), intent(in) :: info
      character(len=:), allocatable :: text
      integer :: i
      text = ''
      do i = 1, size(info%label)
         if (info%label(i) == c_null_char) exit
         text = text//info%label(i)
      end do
   end function info_label

   function c_text(p) result(text)
      type(c_ptr), intent(in) :: p
      character(len=:), allocatable :: text
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      text = ''
      if (.not. c_associated(p)) return
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate (character(len=n) :: text)
      do i = 1, n
         text(i:i) = chars(i)
      end do
   end function c_text

end module sqc_hip_elements
