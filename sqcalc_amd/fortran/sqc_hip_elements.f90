!> Scattering data for the species labels of -m TYPE:LABEL.
!!
!! Neutron bound coherent lengths b [fm] (Neutron News 3 (1992) 29-37) per
!! element and IT92 X-ray form-factor coefficients (International Tables C
!! (1992) 6.1.1.4, f(q) = c + sum_i a_i exp(-b_i (q/4 pi)^2)) per label: the
!! published values the reference tabulates in src/sqc_element_data.f90 and
!! looks up in src/sqc_elements.f90:181-239.  An ion or valence label (O2-,
!! Si4+) takes the neutron length of its element.  The same table backs the
!! Python host (sqcalc_amd/elements.py); a test compares the two.
module sqc_hip_elements
   use, intrinsic :: iso_fortran_env, only: real64
   implicit none
   private
   public :: element_of, neutron_b, xray_coefficients, n_neutron, n_xray, neutron_label, &
             xray_label

   integer, parameter :: rk = real64
   integer, parameter :: n_neutron = 9, n_xray = 15

   character(len=2), parameter :: nb_sym(n_neutron) = &
      [character(len=2) :: 'H', 'C', 'O', 'Na', 'Al', 'Si', 'Cl', 'Ca', 'Ni']
   real(rk), parameter :: nb_val(n_neutron) = &
      [-3.739_rk, 6.646_rk, 5.803_rk, 3.63_rk, 3.449_rk, 4.1491_rk, 9.577_rk, 4.7_rk, 10.3_rk]

   character(len=4), parameter :: xr_sym(n_xray) = [character(len=4) :: &
      'H', 'C', 'O', 'Na', 'Al', 'Si', 'Cl', 'Ca', 'Ni', 'O2-', 'Si4+', 'Na1+', 'Ca2+', &
      'Ni2+', 'Al3+']
   ! a1..a4, b1..b4, c per label
   real(rk), parameter :: xr_val(9, n_xray) = reshape([ &
      0.493002_rk, 0.322912_rk, 0.140191_rk, 0.04081_rk, 10.5109_rk, 26.1257_rk, 3.14236_rk, &
      57.7997_rk, 0.003038_rk, &
      2.31_rk, 1.02_rk, 1.5886_rk, 0.865_rk, 20.8439_rk, 10.2075_rk, 0.5687_rk, 51.6512_rk, &
      0.2156_rk, &
      3.0485_rk, 2.2868_rk, 1.5463_rk, 0.867_rk, 13.2771_rk, 5.7011_rk, 0.3239_rk, &
      32.9089_rk, 0.2508_rk, &
      4.7626_rk, 3.1736_rk, 1.2674_rk, 1.1128_rk, 3.285_rk, 8.8422_rk, 0.3136_rk, 129.424_rk, &
      0.676_rk, &
      6.4202_rk, 1.9002_rk, 1.5936_rk, 1.9646_rk, 3.0387_rk, 0.7426_rk, 31.5472_rk, &
      85.0886_rk, 1.1151_rk, &
      6.2915_rk, 3.0353_rk, 1.9891_rk, 1.541_rk, 2.4386_rk, 32.3337_rk, 0.6785_rk, &
      81.6937_rk, 1.1407_rk, &
      11.4604_rk, 7.1964_rk, 6.2556_rk, 1.6455_rk, 0.0104_rk, 1.1662_rk, 18.5194_rk, &
      47.7784_rk, -9.5574_rk, &
      8.6266_rk, 7.3873_rk, 1.5899_rk, 1.0211_rk, 10.4421_rk, 0.6599_rk, 85.7484_rk, &
      178.437_rk, 1.3751_rk, &
      12.8376_rk, 7.292_rk, 4.4438_rk, 2.38_rk, 3.8785_rk, 0.2565_rk, 12.1763_rk, 66.3421_rk, &
      1.0341_rk, &
      3.7504_rk, 2.84294_rk, 1.54298_rk, 1.62091_rk, 16.5151_rk, 6.59203_rk, 0.319201_rk, &
      43.3486_rk, 0.24206_rk, &
      4.43918_rk, 3.20345_rk, 1.19453_rk, 0.41653_rk, 1.64167_rk, 3.43757_rk, 0.2149_rk, &
      6.65365_rk, 0.746297_rk, &
      3.2565_rk, 3.9362_rk, 1.3998_rk, 1.0032_rk, 2.6671_rk, 6.1153_rk, 0.2001_rk, 14.039_rk, &
      0.404_rk, &
      15.6348_rk, 7.9518_rk, 8.4372_rk, 0.8537_rk, -0.0074_rk, 0.6089_rk, 10.3116_rk, &
      25.9905_rk, -14.875_rk, &
      11.4166_rk, 7.4005_rk, 5.3442_rk, 0.9773_rk, 3.6766_rk, 0.2449_rk, 8.873_rk, 22.1626_rk, &
      0.8614_rk, &
      4.17448_rk, 3.3876_rk, 1.20296_rk, 0.528137_rk, 1.93816_rk, 4.14553_rk, 0.228753_rk, &
      8.28524_rk, 0.706786_rk], [9, n_xray])

contains

   !> Element symbol of a label: a capital letter and an optional lower-case one.
   function element_of(label) result(sym)
      character(len=*), intent(in) :: label
      character(len=:), allocatable :: sym
      sym = ''
      if (len_trim(label) < 1) return
      if (label(1:1) < 'A' .or. label(1:1) > 'Z') return
      sym = label(1:1)
      if (len_trim(label) >= 2) then
         if (label(2:2) >= 'a' .and. label(2:2) <= 'z') sym = label(1:2)
      end if
   end function element_of

   !> Neutron length of a label's element; ok = .false. if not tabulated.
   subroutine neutron_b(label, b, ok)
      character(len=*), intent(in) :: label
      real(rk), intent(out) :: b
      logical, intent(out) :: ok
      integer :: i
      character(len=:), allocatable :: sym
      sym = element_of(label)
      b = 0.0_rk
      ok = .false.
      do i = 1, n_neutron
         if (trim(nb_sym(i)) == sym) then
            b = nb_val(i)
            ok = .true.
            return
         end if
      end do
   end subroutine neutron_b

   !> IT92 coefficients of a label (exact match); ok = .false. if not tabulated.
   subroutine xray_coefficients(label, coef, ok)
      character(len=*), intent(in) :: label
      real(rk), intent(out) :: coef(9)
      logical, intent(out) :: ok
      integer :: i
      coef = 0.0_rk
      ok = .false.
      do i = 1, n_xray
         if (trim(xr_sym(i)) == trim(label)) then
            coef = xr_val(:, i)
            ok = .true.
            return
         end if
      end do
   end subroutine xray_coefficients

   function neutron_label(i) result(sym)
      integer, intent(in) :: i
      character(len=:), allocatable :: sym
      sym = trim(nb_sym(i))
   end function neutron_label

   function xray_label(i) result(sym)
      integer, intent(in) :: i
      character(len=:), allocatable :: sym
      sym = trim(xr_sym(i))
   end function xray_label

end module sqc_hip_elements
