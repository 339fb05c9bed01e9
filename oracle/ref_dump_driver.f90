!> Test driver around the reference's own dump reader (oracle only).
!!
!! Built by `make ref` against the reference's src/sqc_kinds.f90,
!! src/sqc_cell.f90 and src/sqc_dump.f90, compiled where they lie.  Reads the
!! dump named by the first argument with lammps_dump_reader_t and writes, per
!! frame, a raw stream record to the second argument:
!!   int32 ierr, natoms, timestep; float64 cell(9), origin(3); int32 pbc(3);
!!   float64 pos(3, natoms); int32 id(natoms), type_id(natoms); int32 n_ambiguous
!! After a nonzero ierr only that int32 is written and the program stops.
program ref_dump_driver
   use sqc_kinds, only: rk
   use sqc_dump, only: frame_t, lammps_dump_reader_t
   use, intrinsic :: iso_fortran_env, only: int32, real64
   implicit none
   type(lammps_dump_reader_t) :: reader
   type(frame_t) :: frame
   character(len=4096) :: inp, outp
   integer :: ierr, u, k
   integer(int32) :: pbc(3)

   call get_command_argument(1, inp)
   call get_command_argument(2, outp)
   open (newunit=u, file=trim(outp), access='stream', form='unformatted', status='replace')
   call reader%open(trim(inp), ierr)
   if (ierr /= 0) then
      write (u) int(ierr, int32)
      close (u)
      stop
   end if
   do
      call reader%next_frame(frame, ierr)
      if (ierr /= 0) then
         write (u) int(ierr, int32)
         exit
      end if
      do k = 1, 3
         pbc(k) = merge(1_int32, 0_int32, frame%pbc(k))
      end do
      write (u) 0_int32, int(frame%natoms, int32), int(frame%timestep, int32)
      write (u) real(frame%cell%a, real64), real(frame%cell%origin, real64), pbc
      write (u) real(frame%pos(:, 1:frame%natoms), real64)
      write (u) int(frame%id(1:frame%natoms), int32), int(frame%type_id(1:frame%natoms), int32)
      write (u) int(reader%n_ambiguous, int32)
   end do
   close (u)
   call reader%close()
end program ref_dump_driver
