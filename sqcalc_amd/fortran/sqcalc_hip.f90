!> sqcalc_hip static: the reciprocal-space S(q) of a LAMMPS trajectory on one
!! HIP device, with the options, exit codes and `# q S(q) S(a-b)...` table of
!! `sqcalc static --device gpu` (app/main.f90, src/sqc_options.f90:180-700,
!! src/sqc_structure_factor.f90:771-827).  Every per-frame step runs in
!! libsqc_hip; this program reads frames, owns the options and writes tables.
!!
!!   sqcalc_hip static -i traj.dump [-m 1:Si,2:O] [-w unit|neutron|xray]
!!                     [--qmin Q] [--qmax Q] [--nq N] [--eps E]
!!                     [--precision double|single] [--norm mean|self|n]
!!                     [--partials|--no-partials] [-fz] [--gpu-id K]
!!                     [--grid FILE] [--xrd FILE --xrd-lambda L [--xrd-range MIN MAX]
!!                     [--xrd-step S] [--lp|--no-lp]] [--tune | --tiles X,Y,C,L] [--quiet]
!!                     [OUT|-]
program sqcalc_hip
   use, intrinsic :: iso_c_binding
   use, intrinsic :: iso_fortran_env, only: error_unit, output_unit, real64, int64
   use sqc_hip
   use sqc_hip_elements
   use sqc_hip_dump
   implicit none

   integer, parameter :: rk = real64, ik = int64
   integer, parameter :: batch = 8
   character(len=*), parameter :: version = 'sqcalc_hip 1.0 (libsqc_hip ABI 1)'
   real(rk), parameter :: deg = acos(-1.0_rk)/180.0_rk

   ! options (defaults of src/sqc_options.f90)
   character(len=:), allocatable :: input, output, grid_output, mapping_text, msg
   integer :: weight_kind = 0, norm = 1, precision = 0, gpu_id = 0, nq = 500
   real(rk) :: qmin = 0.0_rk, qmax = 20.0_rk, eps = 1.0e-9_rk
   logical :: eps_given = .false., partials = .true., faber_ziman = .false., quiet = .false.
   logical :: tune = .false., want_grid = .false.
   ! --tiles X,Y,C,L: pinned spread tile, y-pass chunk and x-pass line group
   integer(c_int) :: tiles(4) = 0
   logical :: qmin_given = .false., qmax_given = .false., weight_given = .false.
   ! powder XRD (src/sqc_options.f90:90-104, :620-660)
   character(len=:), allocatable :: xrd_output
   real(rk) :: xrd_lambda = 0.0_rk, xrd_min = 1.0_rk, xrd_max = 179.0_rk, xrd_step = 0.0_rk
   logical :: xrd_lp = .true., xrd_opts = .false.

   ! mapping -m TYPE:LABEL
   integer(ik), allocatable :: map_type(:)
   character(len=8), allocatable :: map_label(:)

   type(dump_reader_t) :: reader
   type(dump_frame_t) :: frame
   type(sqc_hip_config_t) :: cfg
   type(sqc_hip_info_t) :: info
   type(c_ptr) :: plan
   integer(c_int) :: rc
   integer :: ierr, nspec, a, b, unit_out, s
   integer(ik) :: natoms, j, nb, nframes, got
   integer :: cur
   logical :: more
   integer(ik), allocatable :: spec_type(:), counts(:)
   integer(c_int32_t), allocatable, target :: species_of(:)
   real(c_double), allocatable, target :: amp(:), xcoef(:, :), pos(:, :, :, :)
   real(c_double), allocatable, target :: num(:), part(:, :, :), den(:), gnum(:)
   integer(c_int64_t), allocatable :: shell_count(:)
   integer(c_int64_t) :: nfr
   real(rk) :: ref_cell(3, 3)
   character(len=8), allocatable :: spec_label(:)

   call parse_options()

   if (.not. sqc_hip_layout_ok()) then
      write (error_unit, '(a)') 'sqcalc_hip: libsqc_hip uses an incompatible sqc_hip_config '// &
         'layout; rebuild the library and this program from the same include/sqc_hip.h'
      stop 3
   end if

   ! --- first frame -------------------------------------------------------
   ! one parse thread fewer than the host's: the thread that waits on the
   ! device spins, and a parse worker sharing its core slows every frame
   call reader%open(input, ierr, -1)
   if (ierr /= 0) then
      write (error_unit, '(a)') 'sqcalc_hip: cannot open dump file "'//input//'"'
      stop 4
   end if
   call reader%next_frame(frame, ierr)
   if (ierr /= 0) then
      write (error_unit, '(a,i0,a)') 'sqcalc_hip: no complete frame found in "'//input// &
         '" (reader error ', ierr, ')'
      stop 5
   end if
   natoms = frame%natoms
   ref_cell = frame%cell

   ! --- species: the type ids present, ascending (src/sqc_structure_factor.f90:1127-1144)
   call classify()
   call species_weights()

   ! --- plan ------------------------------------------------------------------
   cfg%device = int(gpu_id, c_int32_t)
   cfg%precision = int(precision, c_int32_t)
   cfg%weight_kind = int(weight_kind, c_int32_t)
   cfg%norm = int(norm, c_int32_t)
   cfg%partials = merge(1_c_int32_t, 0_c_int32_t, partials)
   cfg%want_grid = merge(1_c_int32_t, 0_c_int32_t, want_grid .or. allocated(xrd_output))
   cfg%mode_partials = merge(1_c_int32_t, 0_c_int32_t, partials .and. allocated(xrd_output))
   cfg%nq = int(nq, c_int32_t)
   cfg%qmin = qmin
   cfg%qmax = qmax
   ! single precision without --eps runs at 1e-5, as the reference does
   ! (app/main.f90:224-230)
   cfg%eps = eps
   if (precision == sqc_hip_f32 .and. .not. eps_given) cfg%eps = 1.0e-5_rk
   cfg%cell = reshape(frame%cell, [9])
   cfg%origin = frame%origin
   cfg%natoms = natoms
   cfg%nspecies = int(nspec, c_int32_t)
   cfg%species_of = c_loc(species_of)
   cfg%amp_const = c_loc(amp)
   cfg%xray_coef = c_loc(xcoef)
   cfg%tune = merge(1_c_int32_t, 0_c_int32_t, tune)
   cfg%max_batch = batch
   rc = sqc_hip_plan_create(plan, cfg)
   if (rc /= sqc_hip_ok .and. rc /= sqc_hip_warn_eps_too_small) then
      write (error_unit, '(a)') 'sqcalc_hip: '//sqc_hip_message(rc)//': '//sqc_hip_detail()
      stop 6
   end if
   if (tiles(1) > 0) then
      rc = sqc_hip_set_tiles(plan, tiles(1), tiles(2), tiles(3), tiles(4))
      if (rc /= sqc_hip_ok) then
         write (error_unit, '(a)') 'sqcalc_hip: '//sqc_hip_message(rc)//': '//sqc_hip_detail()
         stop 6
      end if
   end if
   rc = sqc_hip_plan_info(plan, info)
   if (.not. quiet) call report()

   ! --- frames: the first one, then whole batches straight from the reader.
   ! Two buffers: the library parses batch k + 1 into one on a thread of its
   ! own while batch k runs through the device from the other.
   allocate (pos(3, natoms, batch, 2))
   cur = 1
   pos(:, :, 1, cur) = frame%pos
   call reader%read_frames(pos(:, :, 2:batch, cur), natoms, ref_cell, got, ierr)
   if (ierr /= 0) call reader_failure(ierr)
   nb = 1 + got
   nframes = 0
   do
      more = nb == batch
      if (more) then
         call reader%read_begin(c_loc(pos(1, 1, 1, 3 - cur)), int(batch, ik), natoms, ref_cell, ierr)
         if (ierr /= 0) call reader_failure(ierr)
      end if
      call flush_batch()
      if (.not. more) exit
      call reader%read_end(got, ierr)
      if (ierr /= 0) call reader_failure(ierr)
      cur = 3 - cur
      nb = got
      if (nb == 0) exit
   end do
   call reader%close()

   ! --- results -----------------------------------------------------------------
   allocate (num(nq), part(nq, nspec, nspec), den(nq), shell_count(nq))
   if (want_grid .or. allocated(xrd_output)) then
      allocate (gnum(info%nmodes))
      rc = sqc_hip_fetch(plan, num, c_loc(part), c_loc(gnum), nfr)
   else
      rc = sqc_hip_fetch(plan, num, c_loc(part), c_null_ptr, nfr)
   end if
   if (rc == sqc_hip_ok) rc = sqc_hip_shells(plan, den, shell_count)
   if (rc /= sqc_hip_ok) then
      write (error_unit, '(a)') 'sqcalc_hip: '//sqc_hip_message(rc)//': '//sqc_hip_detail()
      stop 19
   end if
   if (.not. quiet) write (error_unit, '(a,i0,a)') '  frames     : ', nfr, ' accumulated'

   if (output == '-') then
      unit_out = output_unit
   else
      open (newunit=unit_out, file=output, status='replace', action='write', iostat=ierr)
      if (ierr /= 0) then
         write (error_unit, '(a)') 'sqcalc_hip: cannot open output file "'//output//'"'
         stop 11
      end if
   end if
   call write_table(unit_out)
   if (unit_out /= output_unit) close (unit_out)
   if (want_grid) call write_grid()
   if (allocated(xrd_output)) call write_xrd()
   rc = sqc_hip_destroy(plan)

contains

   !> Exit codes of the reference driver for reader failures (app/main.f90:225-247).
   subroutine reader_failure(code)
      integer, intent(in) :: code
      select case (code)
      case (10)
         write (error_unit, '(a)') 'sqcalc_hip: the atom order (id) changes between frames; '// &
            'sort the dump (for example dump_modify ... sort id) or keep the order fixed'
         stop 24
      case (11)
         write (error_unit, '(a)') 'sqcalc_hip: the number of atoms changes between frames'
         stop 25
      case (12)
         write (error_unit, '(a)') 'sqcalc_hip: number of atoms changes between frames'
         stop 8
      case (13)
         write (error_unit, '(a)') 'sqcalc_hip: the simulation box changes between frames'
         stop 9
      case default
         write (error_unit, '(a,i0,a,i0,a)') 'sqcalc_hip: malformed dump near line ', &
            reader%lineno, ' (reader error ', code, '): '//reader%message()
         stop 7
      end select
   end subroutine reader_failure

   subroutine flush_batch()
      if (nb == 0) return
      rc = sqc_hip_accumulate(plan, pos(:, :, :, cur), int(nb, c_int64_t))
      if (rc /= sqc_hip_ok) then
         write (error_unit, '(a)') 'sqcalc_hip: '//sqc_hip_message(rc)//': '//sqc_hip_detail()
         stop 10
      end if
      nframes = nframes + nb
      nb = 0
   end subroutine flush_batch

   subroutine classify()
      integer(ik) :: t, tmax
      logical, allocatable :: present(:)
      integer :: k
      tmax = maxval(frame%type_id)
      if (minval(frame%type_id) < 1) then
         write (error_unit, '(a)') 'sqcalc_hip: atom types must be positive integers'
         stop 5
      end if
      allocate (present(tmax))
      present = .false.
      do j = 1, natoms
         present(frame%type_id(j)) = .true.
      end do
      nspec = count(present)
      if (nspec > 32) then
         write (error_unit, '(a)') 'sqcalc_hip: at most 32 atom types are supported'
         stop 6
      end if
      allocate (spec_type(nspec), counts(nspec), spec_label(nspec))
      k = 0
      do t = 1, tmax
         if (present(t)) then
            k = k + 1
            spec_type(k) = t
         end if
      end do
      allocate (species_of(natoms))
      counts = 0
      do j = 1, natoms
         do k = 1, nspec
            if (spec_type(k) == frame%type_id(j)) exit
         end do
         species_of(j) = int(k - 1, c_int32_t)
         counts(k) = counts(k) + 1
      end do
      do k = 1, nspec
         spec_label(k) = ' '
         if (allocated(map_type)) then
            do a = 1, size(map_type)
               if (map_type(a) == spec_type(k)) spec_label(k) = map_label(a)
            end do
         end if
      end do
   end subroutine classify

   !> Amplitudes of the species present (validate_mapping and
   !! weight_scheme_t%amplitude, app/main.f90:515-551, src/sqc_weights.f90:136-154).
   subroutine species_weights()
      integer :: k, ierr
      character(len=:), allocatable :: message
      type(sqc_hip_species_info_t) :: sinfo
      allocate (amp(nspec), xcoef(9, nspec))
      amp = 1.0_rk
      xcoef = 0.0_rk
      do k = 1, nspec
         if (weight_kind == sqc_hip_weight_unit) cycle
         if (len_trim(spec_label(k)) == 0) then
            write (error_unit, '(a,i0,a)') 'sqcalc_hip: atom type ', spec_type(k), &
               ' appears in the dump but is missing from the element mapping'
            stop 14
         end if
         call resolve_species(spec_label(k), sinfo, ierr, message)
         if (weight_kind == sqc_hip_weight_neutron) then
            if (sinfo%has_neutron == 0) then
               write (error_unit, '(a)') 'sqcalc_hip: no neutron scattering length for '// &
                  trim(spec_label(k))//'; use --weight xray or unit'
               stop 15
            end if
            amp(k) = sinfo%neutron_b
         else
            if (sinfo%has_xray == 0) then
               write (error_unit, '(a)') 'sqcalc_hip: no X-ray form factor for '// &
                  trim(spec_label(k))//'; use --weight neutron or unit'
               stop 15
            end if
            xcoef(:, k) = sinfo%xray_coef
         end if
      end do
   end subroutine species_weights

   subroutine report()
      write (error_unit, '(a,i0,a,i0)') '  atoms      : ', natoms, '   species: ', nspec
      write (error_unit, '(a,3(i0,1x),a,i0)') '  modes      : ', info%modes, '  kept: ', &
         info%nmodes
      write (error_unit, '(a,3(i0,1x),a,i0,a,f5.2)') '  fine grid  : ', info%fine, &
         '  ns: ', info%nspread, '  sigma: ', info%upsampfac
      write (error_unit, '(a,f0.3,a)') '  device mem : ', real(info%device_bytes, rk)/1.0e9_rk, &
         ' GB'
   end subroutine report

   function column_label(a, b) result(label)
      integer, intent(in) :: a, b
      character(len=24) :: label
      character(len=18) :: pair
      if (len_trim(spec_label(a)) > 0 .and. len_trim(spec_label(b)) > 0) then
         pair = trim(spec_label(a))//'-'//trim(spec_label(b))
      else
         write (pair, '(i0,a,i0)') spec_type(a), '-', spec_type(b)
      end if
      if (faber_ziman) then
         label = 'A('//trim(pair)//')'
      else
         label = 'S('//trim(pair)//')'
      end if
   end function column_label

   !> The table of sf_write_results (src/sqc_structure_factor.f90:771-827)
   !! with sf_shell_value / sf_partial_value (:657-719).
   subroutine write_table(u)
      integer, intent(in) :: u
      real(rk) :: frames, dq, qc, v, xa, xb
      frames = real(max(nfr, 1_c_int64_t), rk)
      dq = (qmax - qmin)/real(nq, rk)
      if (partials) then
         write (u, '(a)', advance='no') '# q S(q)'
         do a = 1, nspec
            do b = a, nspec
               write (u, '(a)', advance='no') ' '//column_label(a, b)
            end do
         end do
         write (u, '(a)') ''
      else
         write (u, '(a)') '# q S(q)'
      end if
      do s = 1, nq
         qc = qmin + (real(s, rk) - 0.5_rk)*dq
         v = 0.0_rk
         if (den(s) > 0.0_rk) v = num(s)/(frames*den(s))
         if (partials) then
            write (u, '(f14.6,2x,es20.12)', advance='no') qc, v
            do a = 1, nspec
               do b = a, nspec
                  ! part(s, b, a) holds the C array partial[a][b][s]
                  v = part(s, b, a)/(frames*real(natoms, rk)*real(max(shell_count(s), &
                                                                     1_c_int64_t), rk))
                  if (faber_ziman) then
                     xa = real(counts(a), rk)/real(natoms, rk)
                     xb = real(counts(b), rk)/real(natoms, rk)
                     if (a == b) v = v - xa
                     v = v/(xa*xb) + 1.0_rk
                  end if
                  write (u, '(2x,es20.12)', advance='no') v
               end do
            end do
            write (u, '(a)') ''
         else
            write (u, '(f14.6,2x,es20.12)') qc, v
         end if
      end do
   end subroutine write_table

   !> Per-mode table (src/sqc_structure_factor.f90:815-823): q vector and
   !! S(q) of every kept mode, normalised by the mode's own denominator.
   subroutine write_grid()
      real(c_double), allocatable, target :: qvec(:, :)
      integer(c_int32_t), allocatable, target :: hkl(:, :)
      real(rk) :: frames, q, st, w, sw, sw2, d, v
      integer(ik) :: k
      integer :: u, i, c
      allocate (qvec(3, info%nmodes), hkl(3, info%nmodes))
      rc = sqc_hip_modes(plan, c_loc(hkl), c_loc(qvec), c_null_ptr)
      open (newunit=u, file=grid_output, status='replace', action='write', iostat=ierr)
      if (ierr /= 0 .or. rc /= sqc_hip_ok) then
         write (error_unit, '(a)') 'sqcalc_hip: cannot write the grid file "'//grid_output//'"'
         stop 11
      end if
      frames = real(max(nfr, 1_c_int64_t), rk)
      write (u, '(a)') '# qx qy qz S(q)'
      do k = 1, info%nmodes
         q = norm2(qvec(:, k))
         ! mode_denominator (src/sqc_structure_factor.f90:425-451)
         sw = 0.0_rk
         sw2 = 0.0_rk
         do c = 1, nspec
            if (weight_kind == sqc_hip_weight_xray) then
               st = (q/(4.0_rk*acos(-1.0_rk)))**2
               w = xcoef(9, c)
               do i = 1, 4
                  w = w + xcoef(i, c)*exp(-xcoef(4 + i, c)*st)
               end do
            else
               w = amp(c)
            end if
            sw = sw + real(counts(c), rk)*w
            sw2 = sw2 + real(counts(c), rk)*w*w
         end do
         select case (norm)
         case (sqc_hip_norm_mean)
            d = sw*sw/real(natoms, rk)
         case (sqc_hip_norm_self)
            d = sw2
         case default
            d = real(natoms, rk)
         end select
         v = 0.0_rk
         if (d > 0.0_rk) v = gnum(k)/(frames*d)
         write (u, '(3(f14.8,2x),es20.12)') qvec(1, k), qvec(2, k), qvec(3, k), v
      end do
      close (u)
   end subroutine write_grid

   !> Powder pattern (src/sqc_structure_factor.f90:846-904 bins, :964-1024
   !! table) from the device's per-mode frame sums: gnum for the total and,
   !! with partials, the per-mode Re(rho_a rho_b*) sums for the pair columns.
   subroutine write_xrd()
      real(c_double), allocatable, target :: qvec(:, :), gpart(:, :)
      integer(c_int32_t), allocatable, target :: hkl(:, :)
      real(rk), allocatable :: tot(:), pr(:, :), wsp(:)
      integer, allocatable :: bin(:)
      real(rk) :: frames, q, tth, st, lp, widths(3), cr(3), vol, bound, theta_max, step, fn
      integer(ik) :: k
      integer :: u, nb, i, c, pidx, npair
      logical :: has_pairs
      allocate (qvec(3, info%nmodes), hkl(3, info%nmodes))
      rc = sqc_hip_modes(plan, c_loc(hkl), c_loc(qvec), c_null_ptr)
      has_pairs = partials
      npair = nspec*(nspec + 1)/2
      if (has_pairs .and. nspec > 1) then
         allocate (gpart(info%nmodes, npair))
         if (rc == sqc_hip_ok) rc = sqc_hip_fetch_mode_partials(plan, c_loc(gpart))
      end if
      if (rc /= sqc_hip_ok) then
         write (error_unit, '(a)') 'sqcalc_hip: '//sqc_hip_message(rc)//': '//sqc_hip_detail()
         stop 19
      end if
      ! bin width from the narrowest cell width when --xrd-step is not given
      step = xrd_step
      if (step <= 0.0_rk) then
         vol = abs(ref_cell(1, 1)*(ref_cell(2, 2)*ref_cell(3, 3) - ref_cell(3, 2)*ref_cell(2, 3)) &
                   - ref_cell(1, 2)*(ref_cell(2, 1)*ref_cell(3, 3) - ref_cell(3, 1)*ref_cell(2, 3)) &
                   + ref_cell(1, 3)*(ref_cell(2, 1)*ref_cell(3, 2) - ref_cell(3, 1)*ref_cell(2, 2)))
         do i = 1, 3
            associate (aj => ref_cell(:, mod(i, 3) + 1), ak => ref_cell(:, mod(i + 1, 3) + 1))
               cr = [aj(2)*ak(3) - aj(3)*ak(2), aj(3)*ak(1) - aj(1)*ak(3), aj(1)*ak(2) - aj(2)*ak(1)]
            end associate
            widths(i) = vol/sqrt(sum(cr*cr))
         end do
         ! periodic directions only; none periodic: the first frame's extent
         bound = huge(1.0_rk)
         do i = 1, 3
            if (frame%pbc(i)) bound = min(bound, widths(i))
         end do
         if (bound >= huge(1.0_rk)) then
            do i = 1, 3
               q = maxval(frame%pos(i, :)) - minval(frame%pos(i, :))
               if (q > 0.0_rk) bound = min(bound, q)
            end do
         end if
         if (bound <= 0.0_rk .or. bound >= huge(1.0_rk)) then
            write (error_unit, '(a)') 'sqcalc_hip: cannot derive the XRD bin width from a '// &
               'degenerate box; pass --xrd-step'
            stop 19
         end if
         theta_max = min(0.5_rk*deg*xrd_max, 60.0_rk*deg)
         step = (xrd_lambda/(bound*cos(theta_max)))/deg
      end if
      nb = max(1, ceiling((xrd_max - xrd_min)/step))
      step = (xrd_max - xrd_min)/real(nb, rk)
      allocate (tot(nb), pr(nb, npair), bin(info%nmodes), wsp(nspec))
      tot = 0.0_rk
      pr = 0.0_rk
      do k = 1, info%nmodes
         q = norm2(qvec(:, k))
         tth = 2.0_rk*asin(max(-1.0_rk, min(1.0_rk, q*xrd_lambda/(4.0_rk*acos(-1.0_rk)))))/deg
         i = min(max(int((tth - xrd_min)/step) + 1, 1), nb)
         lp = 1.0_rk
         if (xrd_lp) lp = (1.0_rk + cos(tth*deg)**2)/(sin(0.5_rk*tth*deg)**2*cos(0.5_rk*tth*deg))
         tot(i) = tot(i) + gnum(k)*lp
         if (.not. has_pairs) cycle
         do c = 1, nspec
            if (weight_kind == sqc_hip_weight_xray) then
               st = (q/(4.0_rk*acos(-1.0_rk)))**2
               wsp(c) = xcoef(9, c) + sum(xcoef(1:4, c)*exp(-xcoef(5:8, c)*st))
            else
               wsp(c) = amp(c)
            end if
         end do
         if (nspec == 1) then
            pr(i, 1) = pr(i, 1) + gnum(k)*lp
         else
            pidx = 0
            do a = 1, nspec
               do b = a, nspec
                  pidx = pidx + 1
                  pr(i, pidx) = pr(i, pidx) + wsp(a)*wsp(b)*gpart(k, pidx)*lp
               end do
            end do
         end if
      end do
      open (newunit=u, file=xrd_output, status='replace', action='write', iostat=ierr)
      if (ierr /= 0) then
         write (error_unit, '(a)') 'sqcalc_hip: cannot write the XRD file "'//xrd_output//'"'
         stop 11
      end if
      frames = real(max(nfr, 1_c_int64_t), rk)
      fn = 1.0_rk/(frames*real(natoms, rk))
      if (nfr == 1) then
         write (u, '(a,f0.6,a,f0.4,a,f0.4,a,f0.5,a,i0,a)') '# xrd: lambda ', xrd_lambda, &
            ' A, 2theta ', xrd_min, ' to ', xrd_max, ' deg, step ', step, ' deg, ', nb, &
            ' bins, 1 frame'
      else
         write (u, '(a,f0.6,a,f0.4,a,f0.4,a,f0.5,a,i0,a,i0,a)') '# xrd: lambda ', xrd_lambda, &
            ' A, 2theta ', xrd_min, ' to ', xrd_max, ' deg, step ', step, ' deg, ', nb, &
            ' bins, ', nfr, ' frames'
      end if
      if (xrd_lp) then
         write (u, '(a)') '# I(2theta) = sum over the reciprocal lattice points in each '// &
            'bin of |rho(q)|^2 times the Lorentz-polarization factor, per atom'
      else
         write (u, '(a)') '# I(2theta) = sum over the reciprocal lattice points in each '// &
            'bin of |rho(q)|^2, without the Lorentz-polarization factor, per atom'
      end if
      if (has_pairs) write (u, '(a)') '# the pair columns add up to the total: '// &
         'I = sum_a I(a-a) + 2 sum_{a<b} I(a-b)'
      write (u, '(a)', advance='no') '# 2theta[deg] I'
      if (has_pairs) then
         do a = 1, nspec
            do b = a, nspec
               write (u, '(a)', advance='no') ' I('//pair_name(a, b)//')'
            end do
         end do
      end if
      write (u, '(a)') ''
      do i = 1, nb
         write (u, '(f12.4,2x,es20.12)', advance='no') xrd_min + (real(i, rk) - 0.5_rk)*step, &
            tot(i)*fn
         if (has_pairs) then
            do c = 1, npair
               write (u, '(2x,es20.12)', advance='no') pr(i, c)*fn
            end do
         end if
         write (u, '(a)') ''
      end do
      close (u)
   end subroutine write_xrd

   function pair_name(a, b) result(pair)
      integer, intent(in) :: a, b
      character(len=:), allocatable :: pair
      character(len=18) :: tmp
      if (len_trim(spec_label(a)) > 0 .and. len_trim(spec_label(b)) > 0) then
         pair = trim(spec_label(a))//'-'//trim(spec_label(b))
      else
         write (tmp, '(i0,a,i0)') spec_type(a), '-', spec_type(b)
         pair = trim(tmp)
      end if
   end function pair_name

   !> The scattering tables this program weights with, one line per entry:
   !! `neutron LABEL b` and `xray LABEL a1..a4 b1..b4 c`.
   subroutine print_elements()
      type(sqc_hip_species_info_t) :: sinfo
      integer :: k
      do k = 0, species_table_size(0) - 1
         call species_table_row(0, k, sinfo)
         if (sinfo%has_neutron == 0) cycle
         write (output_unit, '(a,1x,a,1x,es24.16)') 'neutron', info_label(sinfo), sinfo%neutron_b
      end do
      do k = 0, species_table_size(1) - 1
         call species_table_row(1, k, sinfo)
         write (output_unit, '(a,1x,a,9(1x,es24.16))') 'xray', info_label(sinfo), sinfo%xray_coef
      end do
   end subroutine print_elements

   subroutine usage(u)
      integer, intent(in) :: u
      write (u, '(a)') 'usage: sqcalc_hip static -i DUMP [-m 1:Si,2:O] [-w unit|neutron|xray]'
      write (u, '(a)') '         [--qmin Q] [--qmax Q] [--nq N] [--eps E] [--precision double|single]'
      write (u, '(a)') '         [--norm mean|self|n] [--partials|--no-partials] [-fz] [--gpu-id K]'
      write (u, '(a)') '         [--device gpu] [--grid FILE] [--tune | --tiles X,Y,C,L] [--quiet] [OUT|-]'
      write (u, '(a)') '         [--xrd FILE --xrd-lambda L [--xrd-range MIN MAX] [--xrd-step S]'
      write (u, '(a)') '          [--lp|--no-lp]]'
   end subroutine usage

   subroutine fail_usage(text)
      character(len=*), intent(in) :: text
      write (error_unit, '(a)') 'sqcalc_hip: '//text
      call usage(error_unit)
      stop 2
   end subroutine fail_usage

   function lower(text) result(out)
      character(len=*), intent(in) :: text
      character(len=len(text)) :: out
      integer :: i, c
      out = text
      do i = 1, len_trim(out)
         c = iachar(out(i:i))
         if (c >= iachar('A') .and. c <= iachar('Z')) out(i:i) = achar(c + 32)
      end do
   end function lower

   subroutine parse_mapping(text)
      character(len=*), intent(in) :: text
      integer :: p, q, colon, n, ios, ierr
      character(len=:), allocatable :: item, message
      type(sqc_hip_species_info_t) :: sinfo
      n = 1
      do p = 1, len_trim(text)
         if (text(p:p) == ',') n = n + 1
      end do
      allocate (map_type(n), map_label(n))
      p = 1
      do a = 1, n
         q = index(text(p:), ',')
         if (q == 0) then
            item = trim(text(p:))
         else
            item = text(p:p + q - 2)
            p = p + q
         end if
         colon = index(item, ':')
         if (colon < 2 .or. colon == len(item)) call fail_usage('bad mapping entry "'//item// &
                                                               '" (use TYPE:SPECIES, e.g. 1:Si)')
         read (item(:colon - 1), *, iostat=ios) map_type(a)
         if (ios /= 0 .or. map_type(a) < 1) call fail_usage('bad type id in mapping "'//item//'"')
         ! every label resolves to a species whatever the weights, and is kept
         ! in its canonical spelling (src/sqc_weights.f90:49-93)
         call resolve_species(item(colon + 1:), sinfo, ierr, message)
         if (ierr /= 0) call fail_usage(message)
         map_label(a) = info_label(sinfo)
      end do
   end subroutine parse_mapping

   subroutine parse_options()
      integer :: i, n, ios, eq
      character(len=1024) :: arg
      character(len=:), allocatable :: name, value
      logical :: inline
      n = command_argument_count()
      if (n < 1) call fail_usage('expected a subcommand: "sqcalc_hip static ..."')
      call get_command_argument(1, arg)
      select case (trim(arg))
      case ('static')
      case ('-h', '--help')
         call usage(output_unit)
         stop
      case ('-v', '--version')
         write (output_unit, '(a)') version
         stop
      case ('elements')
         call print_elements()
         stop
      case ('dyn')
         call fail_usage('only the static subcommand runs on this path')
      case default
         call fail_usage('expected a subcommand as the first argument, got "'//trim(arg)//'"')
      end select
      output = '-'
      i = 2
      do while (i <= n)
         call get_command_argument(i, arg)
         if (len_trim(arg) == 0) then
            i = i + 1
            cycle
         end if
         if (arg(1:1) /= '-' .or. trim(arg) == '-') then
            output = trim(arg)
            i = i + 1
            cycle
         end if
         eq = index(arg, '=')
         inline = eq > 0
         if (inline) then
            name = arg(:eq - 1)
            value = trim(arg(eq + 1:))
         else
            name = trim(arg)
            value = ''
         end if
         select case (name)
         case ('-h', '--help')
            call usage(output_unit)
            stop
         case ('-v', '--version')
            write (output_unit, '(a)') version
            stop
         case ('--quiet')
            quiet = .true.
         case ('--partials')
            partials = .true.
         case ('--no-partials')
            partials = .false.
         case ('-fz', '--faber-ziman')
            faber_ziman = .true.
         case ('--tune')
            tune = .true.
         case ('--lp', '--no-lp')
            xrd_lp = name == '--lp'
            xrd_opts = .true.
         case ('--xrd-range')
            ! the two limits are separate arguments (src/sqc_options.f90:307-313)
            if (.not. inline) then
               i = i + 1
               if (i > n) call fail_usage('--xrd-range takes two values: MIN2TH MAX2TH')
               call get_command_argument(i, arg)
               value = trim(arg)
            end if
            read (value, *, iostat=ios) xrd_min
            i = i + 1
            if (i > n .or. ios /= 0) call fail_usage('--xrd-range takes two values: MIN2TH MAX2TH')
            call get_command_argument(i, arg)
            read (arg, *, iostat=ios) xrd_max
            if (ios /= 0) call fail_usage('--xrd-range takes two numbers: MIN2TH MAX2TH in degrees')
            xrd_opts = .true.
         case ('-i', '--input', '-m', '--mapping', '-w', '--weight', '--qmin', '--qmax', '--nq', &
               '--eps', '--precision', '--device', '--gpu-id', '--norm', '--grid', '--xrd', &
               '--xrd-lambda', '--xrd-step', '--tiles')
            if (.not. inline) then
               i = i + 1
               if (i > n) call fail_usage(name//' needs a value')
               call get_command_argument(i, arg)
               value = trim(arg)
            end if
            select case (name)
            case ('-i', '--input')
               input = value
            case ('-m', '--mapping')
               mapping_text = value
               call parse_mapping(value)
            case ('-w', '--weight')
               weight_given = .true.
               select case (trim(lower(value)))
               case ('unit', 'none', '1')
                  weight_kind = sqc_hip_weight_unit
               case ('neutron', 'n')
                  weight_kind = sqc_hip_weight_neutron
               case ('xray', 'x-ray', 'x_ray', 'x')
                  weight_kind = sqc_hip_weight_xray
               case default
                  call fail_usage('unknown weight scheme "'//value//'" (use unit, neutron or xray)')
               end select
            case ('--qmin')
               read (value, *, iostat=ios) qmin
               if (ios /= 0 .or. qmin < 0.0_rk) call fail_usage('--qmin must be a non-negative number')
               qmin_given = .true.
            case ('--qmax')
               read (value, *, iostat=ios) qmax
               if (ios /= 0 .or. qmax <= 0.0_rk) call fail_usage('--qmax must be a positive number')
               qmax_given = .true.
            case ('--nq')
               read (value, *, iostat=ios) nq
               if (ios /= 0 .or. nq < 1) call fail_usage('--nq must be a positive integer')
            case ('--eps')
               read (value, *, iostat=ios) eps
               if (ios /= 0 .or. eps <= 0.0_rk) call fail_usage('--eps must be a positive number')
               eps_given = .true.
            case ('--precision')
               select case (value)
               case ('double', 'float64', 'f64')
                  precision = sqc_hip_f64
               case ('single', 'float', 'float32', 'f32')
                  precision = sqc_hip_f32
               case default
                  call fail_usage('unknown precision "'//value//'" (use single or double)')
               end select
            case ('--device')
               select case (value)
               case ('gpu', 'cuda', 'hip')
               case ('cpu')
                  call fail_usage('this program runs the GPU path only (--device gpu)')
               case default
                  call fail_usage('unknown device "'//value//'" (use gpu)')
               end select
            case ('--tiles')
               ! spread tile x, y, y-pass column chunk, x-pass lines per group
               read (value, *, iostat=ios) tiles
               if (ios /= 0 .or. any(tiles < 1)) &
                  call fail_usage('--tiles takes four positive integers: X,Y,CHUNK,LINES')
            case ('--gpu-id')
               read (value, *, iostat=ios) gpu_id
               if (ios /= 0 .or. gpu_id < 0) call fail_usage('GPU id must be a non-negative integer')
            case ('--norm')
               select case (value)
               case ('mean', 'fz')
                  norm = sqc_hip_norm_mean
               case ('self')
                  norm = sqc_hip_norm_self
               case ('n', 'natom')
                  norm = sqc_hip_norm_natom
               case default
                  call fail_usage('unknown normalization "'//value//'" (use mean, self or n)')
               end select
            case ('--grid')
               want_grid = .true.
               grid_output = value
            case ('--xrd')
               xrd_output = value
            case ('--xrd-lambda')
               read (value, *, iostat=ios) xrd_lambda
               if (ios /= 0 .or. xrd_lambda <= 0.0_rk) &
                  call fail_usage('--xrd-lambda must be a positive wavelength')
               xrd_opts = .true.
            case ('--xrd-step')
               read (value, *, iostat=ios) xrd_step
               if (ios /= 0 .or. xrd_step <= 0.0_rk) &
                  call fail_usage('--xrd-step must be a positive angle in degrees')
               xrd_opts = .true.
            end select
         case default
            call fail_usage('unknown option "'//name//'"')
         end select
         i = i + 1
      end do
      if (.not. allocated(input)) call fail_usage('missing input dump file (-i)')
      ! --xrd sets the q window and the default weights (src/sqc_options.f90:620-660)
      if (allocated(xrd_output)) then
         if (xrd_lambda <= 0.0_rk) call fail_usage('--xrd needs --xrd-lambda LAMBDA (the incident wavelength)')
         if (qmin_given .or. qmax_given) call fail_usage('--xrd takes the q range from '// &
            '--xrd-lambda and --xrd-range; do not pass --qmin or --qmax with it')
         if (xrd_min <= 0.0_rk .or. xrd_max >= 180.0_rk .or. xrd_max <= xrd_min) &
            call fail_usage('--xrd-range needs 0 < MIN2TH < MAX2TH < 180 degrees')
         if (.not. weight_given) weight_kind = sqc_hip_weight_xray
         qmin = 4.0_rk*acos(-1.0_rk)*sin(0.5_rk*deg*xrd_min)/xrd_lambda
         qmax = 4.0_rk*acos(-1.0_rk)*sin(0.5_rk*deg*xrd_max)/xrd_lambda
      else if (xrd_opts) then
         call fail_usage('--xrd-lambda, --xrd-range, --xrd-step and --lp/--no-lp belong to --xrd FILE')
      end if
      if (qmax <= qmin) call fail_usage('qmax must be larger than qmin')
      if (weight_kind /= sqc_hip_weight_unit .and. .not. allocated(map_type)) &
         call fail_usage('weighted schemes need an element mapping, e.g. -m 1:Si,2:O')
   end subroutine parse_options

end program sqcalc_hip
