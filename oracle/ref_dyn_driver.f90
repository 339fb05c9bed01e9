!> Test driver around the reference's own separable phase factors (oracle only).
!!
!! Built by `make ref` against the reference's src/sqc_kinds.f90 and
!! src/sqc_phase.f90, compiled where they lie.  It runs the amplitude loop of
!! the reference's dyn_accumulate (src/sqc_dynamics.f90:826-851: per atom the
!! projections phi_k = recip(1,k) r1 + recip(2,k) r2 + recip(3,k) r3 and an
!! optional q-line base, dyn_atom_phase :1094-1115, then
!! rho(m, species) += phase_factor(...)) on one thread, with the reference's
!! phase_tables and phase_factor.  Input (stream, first argument):
!!   int32 natoms, nspecies, nmodes, base_on; float64 recip(3,3), base(3);
!!   float64 pos(3, natoms); int32 species(natoms) (1-based); int32 hkl(3, nmodes)
!! Output (stream, second argument): complex(real64) rho(nmodes, nspecies).
program ref_dyn_driver
   use sqc_kinds, only: rk
   use sqc_phase, only: phase_tables, phase_factor
   use, intrinsic :: iso_c_binding, only: c_double_complex
   use, intrinsic :: iso_fortran_env, only: int32, real64
   implicit none
   character(len=4096) :: inp, outp
   integer(int32) :: natoms, nspecies, nmodes, base_on
   real(real64) :: recip(3, 3), basev(3), phi(3), d
   real(real64), allocatable :: pos(:, :)
   integer(int32), allocatable :: species(:), hkl(:, :)
   complex(c_double_complex), allocatable :: rho(:, :), t(:, :)
   complex(c_double_complex) :: base(3)
   integer :: u, i, im, k, index_max(3)

   call get_command_argument(1, inp)
   call get_command_argument(2, outp)
   open (newunit=u, file=trim(inp), access='stream', form='unformatted', status='old')
   read (u) natoms, nspecies, nmodes, base_on
   read (u) recip, basev
   allocate (pos(3, natoms), species(natoms), hkl(3, nmodes))
   read (u) pos, species, hkl
   close (u)
   do k = 1, 3
      index_max(k) = maxval(abs(hkl(k, :)))
   end do
   allocate (t(3, 0:2*maxval(index_max)), rho(nmodes, nspecies))
   rho = (0.0_rk, 0.0_rk)
   do i = 1, natoms
      do im = 1, 3
         phi(im) = recip(1, im)*pos(1, i) + recip(2, im)*pos(2, i) + recip(3, im)*pos(3, i)
      end do
      if (base_on /= 0) then
         d = basev(1)*pos(1, i) + basev(2)*pos(2, i) + basev(3)*pos(3, i)
         base = (1.0_rk, 0.0_rk)
         base(1) = cmplx(cos(d), sin(d), c_double_complex)
         call phase_tables(t, phi, index_max, base)
      else
         call phase_tables(t, phi, index_max)
      end if
      do im = 1, nmodes
         rho(im, species(i)) = rho(im, species(i)) + phase_factor(t, index_max, hkl(:, im))
      end do
   end do
   open (newunit=u, file=trim(outp), access='stream', form='unformatted', status='replace')
   write (u) rho
   close (u)
end program ref_dyn_driver
