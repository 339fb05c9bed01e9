!> LAMMPS text dump reader for the sqcalc_hip command.
!!
!! Reads the frames the reference reader accepts for `sqcalc static`
!! (src/sqc_dump.f90): ITEM: TIMESTEP / NUMBER OF ATOMS / BOX BOUNDS (with
!! or without the xy xz yz tilt line) / ATOMS with an id, a type and one of
!! the coordinate triples x y z, xu yu zu (Cartesian) or xs ys zs, xsu ysu zsu
!! (scaled).  The cell is the restricted triclinic cell of the bounds
!! (src/sqc_cell.f90:39-63); positions come back Cartesian, in file order.
!! Status: 0 a frame was read, 1 end of file before a frame, >1 malformed.
module sqc_hip_dump
   use, intrinsic :: iso_fortran_env, only: real64, int64, iostat_end
   implicit none
   private
   public :: dump_reader_t, dump_frame_t, cell_from_bounds

   integer, parameter :: rk = real64, ik = int64
   integer, parameter :: line_len = 1024, max_cols = 64

   type :: dump_frame_t
      integer(ik) :: natoms = 0
      integer(ik) :: timestep = 0
      real(rk) :: cell(3, 3) = 0.0_rk     ! lattice vectors in the columns
      real(rk) :: origin(3) = 0.0_rk
      integer(ik), allocatable :: id(:)
      integer(ik), allocatable :: type_id(:)
      real(rk), allocatable :: pos(:, :)  ! (3, natoms) Cartesian
   end type dump_frame_t

   type :: dump_reader_t
      integer :: unit = -1
      integer :: lineno = 0
   contains
      procedure :: open => reader_open
      procedure :: next_frame => reader_next
      procedure :: close => reader_close
   end type dump_reader_t

contains

   subroutine reader_open(self, path, ierr)
      class(dump_reader_t), intent(inout) :: self
      character(len=*), intent(in) :: path
      integer, intent(out) :: ierr
      open (newunit=self%unit, file=path, status='old', action='read', iostat=ierr)
      if (ierr /= 0) self%unit = -1
      self%lineno = 0
   end subroutine reader_open

   subroutine reader_close(self)
      class(dump_reader_t), intent(inout) :: self
      if (self%unit >= 0) close (self%unit)
      self%unit = -1
   end subroutine reader_close

   subroutine get_line(self, line, ierr)
      class(dump_reader_t), intent(inout) :: self
      character(len=*), intent(out) :: line
      integer, intent(out) :: ierr
      read (self%unit, '(a)', iostat=ierr) line
      if (ierr == 0) self%lineno = self%lineno + 1
   end subroutine get_line

   !> Restricted triclinic cell from LAMMPS bounds (src/sqc_cell.f90:39-63).
   subroutine cell_from_bounds(lo, hi, tilt, cell, origin)
      real(rk), intent(in) :: lo(3), hi(3), tilt(3)
      real(rk), intent(out) :: cell(3, 3), origin(3)
      real(rk) :: xy, xz, yz, xlo, xhi, ylo, yhi
      xy = tilt(1)
      xz = tilt(2)
      yz = tilt(3)
      xlo = lo(1) - min(0.0_rk, xy, xz, xy + xz)
      xhi = hi(1) - max(0.0_rk, xy, xz, xy + xz)
      ylo = lo(2) - min(0.0_rk, yz)
      yhi = hi(2) - max(0.0_rk, yz)
      cell = 0.0_rk
      cell(:, 1) = [xhi - xlo, 0.0_rk, 0.0_rk]
      cell(:, 2) = [xy, yhi - ylo, 0.0_rk]
      cell(:, 3) = [xz, yz, hi(3) - lo(3)]
      origin = [xlo, ylo, lo(3)]
   end subroutine cell_from_bounds

   subroutine reader_next(self, fr, ierr)
      class(dump_reader_t), intent(inout) :: self
      type(dump_frame_t), intent(inout) :: fr
      integer, intent(out) :: ierr
      character(len=line_len) :: line
      character(len=32) :: names(max_cols)
      real(rk) :: lo(3), hi(3), tilt(3), vals(max_cols), s(3)
      integer :: i, ncol, cid, ctype, cx, kind, ios
      integer(ik) :: j, n
      logical :: triclinic, have_step, have_n, have_box

      have_step = .false.
      have_n = .false.
      have_box = .false.
      n = 0
      do
         call get_line(self, line, ios)
         if (ios == iostat_end) then
            ierr = 1
            if (have_step .or. have_n .or. have_box) ierr = 2
            return
         end if
         if (ios /= 0) then
            ierr = 2
            return
         end if
         line = adjustl(line)
         if (len_trim(line) == 0) cycle
         if (index(line, 'ITEM: TIMESTEP') == 1) then
            call get_line(self, line, ios)
            read (line, *, iostat=ios) fr%timestep
            if (ios /= 0) then
               ierr = 3
               return
            end if
            have_step = .true.
         else if (index(line, 'ITEM: NUMBER OF ATOMS') == 1) then
            call get_line(self, line, ios)
            read (line, *, iostat=ios) n
            if (ios /= 0 .or. n < 1) then
               ierr = 4
               return
            end if
            have_n = .true.
         else if (index(line, 'ITEM: BOX BOUNDS') == 1) then
            triclinic = index(line, 'xy') > 0
            tilt = 0.0_rk
            do i = 1, 3
               call get_line(self, line, ios)
               if (triclinic) then
                  read (line, *, iostat=ios) lo(i), hi(i), tilt(i)
               else
                  read (line, *, iostat=ios) lo(i), hi(i)
               end if
               if (ios /= 0) then
                  ierr = 5
                  return
               end if
            end do
            call cell_from_bounds(lo, hi, tilt, fr%cell, fr%origin)
            have_box = .true.
         else if (index(line, 'ITEM: ATOMS') == 1) then
            if (.not. (have_n .and. have_box)) then
               ierr = 6
               return
            end if
            call split_words(line(12:), names, ncol)
            cid = find(names, ncol, ['id  '])
            ctype = find(names, ncol, ['type'])
            kind = 0
            cx = find(names, ncol, ['x   '])
            if (cx > 0) kind = 1
            if (cx == 0) then
               cx = find(names, ncol, ['xu  '])
               if (cx > 0) kind = 1
            end if
            if (cx == 0) then
               cx = find(names, ncol, ['xs  '])
               if (cx > 0) kind = 2
            end if
            if (cx == 0) then
               cx = find(names, ncol, ['xsu '])
               if (cx > 0) kind = 2
            end if
            if (cid == 0 .or. ctype == 0 .or. cx == 0) then
               ierr = 7
               return
            end if
            if (cx + 2 > ncol) then
               ierr = 7
               return
            end if
            fr%natoms = n
            if (allocated(fr%pos)) then
               if (size(fr%pos, 2) /= n) deallocate (fr%pos, fr%id, fr%type_id)
            end if
            if (.not. allocated(fr%pos)) allocate (fr%pos(3, n), fr%id(n), fr%type_id(n))
            do j = 1, n
               call get_line(self, line, ios)
               if (ios /= 0) then
                  ierr = 8
                  return
               end if
               read (line, *, iostat=ios) vals(1:ncol)
               if (ios /= 0) then
                  ierr = 9
                  return
               end if
               fr%id(j) = nint(vals(cid), ik)
               fr%type_id(j) = nint(vals(ctype), ik)
               if (kind == 1) then
                  fr%pos(:, j) = vals(cx:cx + 2)
               else
                  s = vals(cx:cx + 2)
                  fr%pos(:, j) = fr%origin + matmul(fr%cell, s)
               end if
            end do
            ierr = 0
            return
         end if
      end do
   end subroutine reader_next

   subroutine split_words(text, words, n)
      character(len=*), intent(in) :: text
      character(len=*), intent(out) :: words(:)
      integer, intent(out) :: n
      integer :: i, start
      n = 0
      i = 1
      do while (i <= len_trim(text))
         if (text(i:i) == ' ' .or. text(i:i) == char(9)) then
            i = i + 1
            cycle
         end if
         start = i
         do while (i <= len_trim(text))
            if (text(i:i) == ' ' .or. text(i:i) == char(9)) exit
            i = i + 1
         end do
         if (n < size(words)) then
            n = n + 1
            words(n) = text(start:i - 1)
         end if
      end do
   end subroutine split_words

   integer function find(words, n, key)
      character(len=*), intent(in) :: words(:)
      integer, intent(in) :: n
      character(len=*), intent(in) :: key(1)
      integer :: i
      find = 0
      do i = 1, n
         if (trim(words(i)) == trim(key(1))) then
            find = i
            return
         end if
      end do
   end function find

end module sqc_hip_dump
