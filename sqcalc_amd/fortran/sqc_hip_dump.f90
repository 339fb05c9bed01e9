!> LAMMPS dump frames for the sqcalc_hip command, through the native reader
!! of libsqc_hip (sqc_dump_* in include/sqc_hip.h).
!!
!! The reader follows the reference's lammps_dump_reader_t (src/sqc_dump.f90):
!! the same header records, the restricted triclinic cell of the bounds
!! (src/sqc_cell.f90:39-63), positions unwrapped across frames unless the dump
!! has xu yu zu, and its ierr values as status codes (0 a frame was read, 1 end
!! of file, 2..11 its errors).  Statuses 12 and 13 are the driver's checks of
!! app/main.f90:240-247 (atom count, box), raised by read_frames.  The scaled
!! columns xs ys zs and xsu ysu zsu are read as well.
module sqc_hip_dump
   use, intrinsic :: iso_c_binding
   use, intrinsic :: iso_fortran_env, only: real64, int64
   implicit none
   private
   public :: dump_reader_t, dump_frame_t

   integer, parameter :: rk = real64, ik = int64

   type, bind(C) :: c_dump_frame
      integer(c_int64_t) :: natoms = 0
      integer(c_int64_t) :: timestep = 0
      real(c_double) :: cell(9) = 0.0_c_double
      real(c_double) :: origin(3) = 0.0_c_double
      integer(c_int32_t) :: pbc(3) = 0
      integer(c_int32_t) :: triclinic = 0
      integer(c_int32_t) :: scaled = 0
      integer(c_int32_t) :: has_id = 0
      integer(c_int32_t) :: unwrapped = 0
      integer(c_int32_t) :: unwrapped_from_dump = 0
      integer(c_int64_t) :: n_ambiguous = 0
      integer(c_int64_t) :: lineno = 0
      integer(c_int64_t) :: frames = 0
   end type c_dump_frame

   type :: dump_frame_t
      integer(ik) :: natoms = 0
      integer(ik) :: timestep = 0
      real(rk) :: cell(3, 3) = 0.0_rk     ! lattice vectors in the columns
      real(rk) :: origin(3) = 0.0_rk
      logical :: pbc(3) = .true.
      logical :: unwrapped_from_dump = .false.
      integer(ik), allocatable :: id(:)
      integer(ik), allocatable :: type_id(:)
      real(rk), allocatable :: pos(:, :)  ! (3, natoms) Cartesian, continuous
   end type dump_frame_t

   type :: dump_reader_t
      type(c_ptr) :: handle = c_null_ptr
      integer(ik) :: lineno = 0
   contains
      procedure :: open => reader_open
      procedure :: next_frame => reader_next
      procedure :: read_frames => reader_read_frames
      procedure :: read_begin => reader_read_begin
      procedure :: read_end => reader_read_end
      procedure :: message => reader_message
      procedure :: close => reader_close
   end type dump_reader_t

   interface
      integer(c_int) function sqc_dump_open(reader, path, nthreads) bind(C, name='sqc_dump_open')
         import :: c_int, c_ptr, c_char
         type(c_ptr), intent(out) :: reader
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: nthreads
      end function sqc_dump_open

      integer(c_int) function sqc_dump_next(reader, frame) bind(C, name='sqc_dump_next')
         import :: c_int, c_ptr, c_dump_frame
         type(c_ptr), value :: reader
         type(c_dump_frame), intent(out) :: frame
      end function sqc_dump_next

      integer(c_int) function sqc_dump_copy(reader, pos, id, type_id) bind(C, name='sqc_dump_copy')
         import :: c_int, c_ptr
         type(c_ptr), value :: reader, pos, id, type_id
      end function sqc_dump_copy

      integer(c_int) function sqc_dump_read_frames(reader, maxframes, natoms, cell, pos, nread, &
                                                   last) bind(C, name='sqc_dump_read_frames')
         import :: c_int, c_ptr, c_int64_t, c_double, c_dump_frame
         type(c_ptr), value :: reader
         integer(c_int64_t), value :: maxframes, natoms
         real(c_double), intent(in) :: cell(9)
         real(c_double) :: pos(*)
         integer(c_int64_t), intent(out) :: nread
         type(c_dump_frame), intent(out) :: last
      end function sqc_dump_read_frames

      integer(c_int) function sqc_dump_read_frames_begin(reader, maxframes, natoms, cell, pos) &
         bind(C, name='sqc_dump_read_frames_begin')
         import :: c_int, c_ptr, c_int64_t, c_double
         type(c_ptr), value :: reader, pos
         integer(c_int64_t), value :: maxframes, natoms
         real(c_double), intent(in) :: cell(9)
      end function sqc_dump_read_frames_begin

      integer(c_int) function sqc_dump_read_frames_end(reader, nread) &
         bind(C, name='sqc_dump_read_frames_end')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: reader
         integer(c_int64_t), intent(out) :: nread
      end function sqc_dump_read_frames_end

      integer(c_int) function sqc_dump_close(reader) bind(C, name='sqc_dump_close')
         import :: c_int, c_ptr
         type(c_ptr), value :: reader
      end function sqc_dump_close

      type(c_ptr) function sqc_dump_message(reader) bind(C, name='sqc_dump_message')
         import :: c_ptr
         type(c_ptr), value :: reader
      end function sqc_dump_message

      integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function c_strlen
   end interface

contains

   !> Open a dump; nthreads <= 0 lets the library pick (OMP_NUM_THREADS).
   subroutine reader_open(self, path, ierr, nthreads)
      class(dump_reader_t), intent(inout) :: self
      character(len=*), intent(in) :: path
      integer, intent(out) :: ierr
      integer, intent(in), optional :: nthreads
      integer(c_int) :: nt
      nt = 0
      if (present(nthreads)) nt = int(nthreads, c_int)
      ierr = int(sqc_dump_open(self%handle, trim(path)//c_null_char, nt))
      self%lineno = 0
   end subroutine reader_open

   subroutine reader_close(self)
      class(dump_reader_t), intent(inout) :: self
      integer(c_int) :: rc
      if (c_associated(self%handle)) rc = sqc_dump_close(self%handle)
      self%handle = c_null_ptr
   end subroutine reader_close

   !> The next frame with its atom arrays.
   subroutine reader_next(self, fr, ierr)
      class(dump_reader_t), intent(inout) :: self
      type(dump_frame_t), intent(inout), target :: fr
      integer, intent(out) :: ierr
      type(c_dump_frame) :: h
      integer(c_int64_t), allocatable, target :: id(:)
      integer(c_int32_t), allocatable, target :: ty(:)
      integer(ik) :: n
      ierr = int(sqc_dump_next(self%handle, h))
      self%lineno = h%lineno
      if (ierr /= 0) return
      n = h%natoms
      fr%natoms = n
      fr%timestep = h%timestep
      fr%cell = reshape(h%cell, [3, 3])
      fr%origin = h%origin
      fr%pbc = h%pbc /= 0
      fr%unwrapped_from_dump = h%unwrapped_from_dump /= 0
      if (allocated(fr%pos)) then
         if (size(fr%pos, 2) /= n) deallocate (fr%pos, fr%id, fr%type_id)
      end if
      if (.not. allocated(fr%pos)) allocate (fr%pos(3, n), fr%id(n), fr%type_id(n))
      allocate (id(n), ty(n))
      ierr = int(sqc_dump_copy(self%handle, c_loc(fr%pos), c_loc(id), c_loc(ty)))
      fr%id = id
      fr%type_id = ty
   end subroutine reader_next

   !> Up to size(pos, 3) frames of natoms atoms straight into pos(3, natoms, :),
   !! each checked against natoms and cell; nread < size(pos, 3) at the end.
   subroutine reader_read_frames(self, pos, natoms, cell, nread, ierr)
      class(dump_reader_t), intent(inout) :: self
      real(c_double), intent(inout), contiguous :: pos(:, :, :)
      integer(ik), intent(in) :: natoms
      real(rk), intent(in) :: cell(3, 3)
      integer(ik), intent(out) :: nread
      integer, intent(out) :: ierr
      type(c_dump_frame) :: h
      integer(c_int64_t) :: got
      ierr = int(sqc_dump_read_frames(self%handle, int(size(pos, 3), c_int64_t), &
                                      int(natoms, c_int64_t), reshape(cell, [9]), pos, got, h))
      nread = got
      if (h%lineno > 0) self%lineno = h%lineno
   end subroutine reader_read_frames

   !> Start reading up to maxframes frames into the memory at pos (3 natoms
   !! doubles a frame) on a thread of the library; pos must stay allocated and
   !! untouched until read_end.  The cell is copied before this returns.
   subroutine reader_read_begin(self, pos, maxframes, natoms, cell, ierr)
      class(dump_reader_t), intent(inout) :: self
      type(c_ptr), intent(in) :: pos
      integer(ik), intent(in) :: maxframes, natoms
      real(rk), intent(in) :: cell(3, 3)
      integer, intent(out) :: ierr
      ierr = int(sqc_dump_read_frames_begin(self%handle, int(maxframes, c_int64_t), &
                                            int(natoms, c_int64_t), reshape(cell, [9]), pos))
   end subroutine reader_read_begin

   !> Wait for the batch of read_begin: its frame count and status.
   subroutine reader_read_end(self, nread, ierr)
      class(dump_reader_t), intent(inout) :: self
      integer(ik), intent(out) :: nread
      integer, intent(out) :: ierr
      integer(c_int64_t) :: got
      ierr = int(sqc_dump_read_frames_end(self%handle, got))
      nread = got
   end subroutine reader_read_end

   !> Detail of the error that ended the reader, with its line.
   function reader_message(self) result(text)
      class(dump_reader_t), intent(in) :: self
      character(len=:), allocatable :: text
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: i, n
      p = sqc_dump_message(self%handle)
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate (character(len=n) :: text)
      do i = 1, n
         text(i:i) = chars(i)
      end do
   end function reader_message

end module sqc_hip_dump
