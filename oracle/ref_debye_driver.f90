!> Test driver around the reference's own cell list (oracle only).
!!
!! Built by `make ref` against the reference's src/sqc_kinds.f90,
!! src/sqc_cell.f90 and src/sqc_neighbour_list.f90, compiled where they lie.
!! Per frame it calls the list's begin_frame (positions, unwrapping, Verlet
!! rebuilds) and then runs the pair loop of the reference's debye_accumulate
!! (src/sqc_debye.f90:208-226: displacement pos_j - pos_i + bins . shift,
!! d2 > rmax^2 skipped, bin = int(sqrt(d2) / dr) + 1) on one thread.
!! Input (stream, first argument):
!!   int32 natoms, ntypes, nframes, pbc(3); float64 a(3,3), origin(3), rmax,
!!   dr, skin; int32 type(natoms) (1-based); float64 pos(3, natoms, nframes)
!! Output (stream, second argument): int32 nbins, nbuilds; int64
!!   npairs of the last list; float64 histogram(ntypes, ntypes, nbins)
!!   summed over the frames.
program ref_debye_driver
   use sqc_kinds, only: rk, ik, lk
   use sqc_cell, only: cell_t
   use sqc_neighbour_list, only: neighbour_list_t
   use, intrinsic :: iso_fortran_env, only: int32, int64, real64, error_unit
   implicit none
   character(len=4096) :: inp, outp
   character(len=512) :: message
   integer(int32) :: natoms, ntypes, nframes, pbc_in(3), nbins
   real(real64) :: a(3, 3), origin(3), rmax, dr, skin, d(3), d2, inv_dr
   integer(int32), allocatable :: types(:)
   real(real64), allocatable :: pos(:, :, :), histogram(:, :, :)
   type(cell_t) :: cell
   type(neighbour_list_t) :: nl
   logical :: pbc(3)
   integer :: u, f, i, j, bin, ierr
   integer(lk) :: k

   call get_command_argument(1, inp)
   call get_command_argument(2, outp)
   open (newunit=u, file=trim(inp), access='stream', form='unformatted', status='old')
   read (u) natoms, ntypes, nframes, pbc_in
   read (u) a, origin, rmax, dr, skin
   allocate (types(natoms), pos(3, natoms, nframes))
   read (u) types, pos
   close (u)
   pbc = pbc_in /= 0
   call cell%set_vectors(a)
   cell%origin = origin
   nbins = max(1, ceiling(rmax/dr))
   allocate (histogram(ntypes, ntypes, nbins))
   histogram = 0.0_rk
   inv_dr = 1.0_rk/dr
   call nl%configure(cell, int(natoms, ik), rmax, skin, pbc, ierr, message)
   if (ierr /= 0) then
      write (error_unit, '(a)') trim(message)
      stop 2
   end if
   do f = 1, nframes
      call nl%begin_frame(pos(:, :, f), ierr, message)
      if (ierr /= 0) then
         write (error_unit, '(a)') trim(message)
         stop 3
      end if
      do i = 1, natoms
         do k = nl%pair_start(i), nl%pair_start(i + 1) - 1
            j = int(nl%pair_atom(k))
            d = nl%pos(:, j) - nl%pos(:, i) + matmul(nl%bins, real(nl%pair_shift(:, k), rk))
            d2 = sum(d*d)
            if (d2 > rmax*rmax) cycle
            bin = int(sqrt(d2)*inv_dr) + 1
            if (bin < 1 .or. bin > nbins) cycle
            histogram(types(i), types(j), bin) = histogram(types(i), types(j), bin) + 1.0_rk
         end do
      end do
   end do
   open (newunit=u, file=trim(outp), access='stream', form='unformatted', status='replace')
   write (u) nbins, int(nl%nbuilds, int32), int(nl%npairs, int64), histogram
   close (u)
end program ref_debye_driver
