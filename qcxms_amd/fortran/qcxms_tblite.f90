!> Drop-in replacement of QCxMS's tblite bridge (src/tblite.f90:3-175) that
!> evaluates GFN2-xTB on an MI355X through libqcxms_gpu.so.
!>
!> The module keeps the name, the public symbols and the argument meaning of
!> the reference module, so src/iniqm.f90 compiles against it unchanged:
!>   get_xtb_egrad(num, xyz, charge, multiplicity, method, etemp, output_file,
!>                 qat, energy, gradient, stat, spec_calc)
!>   gfn1_xtb, gfn2_xtb, ipea1_xtb                     (src/tblite.f90:22-40)
!> Status codes as src/tblite.f90:55-58: 0 success, 5 unknown (here: not
!> provided on the GPU) method, -1 failed calculation.  A HIP runtime or
!> argument error of the library (-2) is reported as -1.
!> output_file receives a short log of the call, as the reference's context
!> unit does.  spec_calc (src/tblite.f90:153-166) writes the MO spectrum of
!> the structure, tmp.mspec and qcxms.Mspec.tbxtb, in the layout of
!> write_qmo (src/mo_energ.f90:7-80), from the orbital energies, occupations
!> and Mulliken MO populations the GPU computes (qcxms_get_xtb_mo_spectrum).
module qcxms_tblite
   use, intrinsic :: iso_c_binding, only : c_int32_t, c_double, c_char, c_ptr, &
      & c_f_pointer, c_associated
   implicit none
   private

   public :: get_xtb_egrad
   public :: gfn1_xtb, gfn2_xtb, ipea1_xtb
   public :: method_selector

   integer, parameter :: wp = selected_real_kind(15)

   !> Enumerated method selector (same ids as the reference)
   type :: method_selector
      integer :: id
   end type method_selector

   type(method_selector), parameter :: gfn1_xtb = method_selector(1)
   type(method_selector), parameter :: gfn2_xtb = method_selector(2)
   type(method_selector), parameter :: ipea1_xtb = method_selector(11)

   integer, parameter :: stat_fatal = -1

   !> Hartree to eV (src/xtb_mctc_convert.f90:27)
   real(wp), parameter :: autoev = 27.21138505_wp

   interface
      !> int qcxms_get_xtb_egrad(int32_t nat, const int32_t *num, const double *xyz,
      !>     int32_t charge, int32_t multiplicity, int32_t method, double etemp,
      !>     double *qat, double *energy, double *gradient, int32_t *stat)
      function c_get_xtb_egrad(nat, num, xyz, charge, multiplicity, method, etemp, &
            & qat, energy, gradient, stat) result(rc) bind(C, name="qcxms_get_xtb_egrad")
         import :: c_int32_t, c_double
         integer(c_int32_t), value :: nat, charge, multiplicity, method
         integer(c_int32_t), intent(in) :: num(*)
         real(c_double), intent(in) :: xyz(*)
         real(c_double), value :: etemp
         real(c_double), intent(out) :: qat(*), energy, gradient(*)
         integer(c_int32_t), intent(out) :: stat
         integer(c_int32_t) :: rc
      end function c_get_xtb_egrad

      !> int qcxms_get_xtb_mo_spectrum(int32_t cap, int32_t *nao, int32_t *ihomo,
      !>     double *emo, double *focc, double *qmo)
      function c_get_xtb_mo_spectrum(cap, nao, ihomo, emo, focc, qmo) result(rc) &
            & bind(C, name="qcxms_get_xtb_mo_spectrum")
         import :: c_int32_t, c_double
         integer(c_int32_t), value :: cap
         integer(c_int32_t), intent(out) :: nao, ihomo
         real(c_double), intent(out) :: emo(*), focc(*), qmo(*)
         integer(c_int32_t) :: rc
      end function c_get_xtb_mo_spectrum

      function c_last_error() result(msg) bind(C, name="qcxms_gpu_last_error")
         import :: c_ptr
         type(c_ptr) :: msg
      end function c_last_error
   end interface

contains

!> Energy, gradient and Mulliken charges of one structure (Bohr, Hartree,
!> Hartree/Bohr, e), fresh wavefunction per call as src/tblite.f90:133.
subroutine get_xtb_egrad(num, xyz, charge, multiplicity, method, etemp, &
      & output_file, qat, energy, gradient, stat, spec_calc)
   integer, intent(in) :: num(:)
   real(wp), intent(in) :: xyz(:, :)
   integer, intent(in) :: charge
   integer, intent(in) :: multiplicity
   type(method_selector), intent(in) :: method
   real(wp), intent(in) :: etemp
   character(len=*), intent(in) :: output_file
   real(wp), intent(out) :: qat(:)
   real(wp), intent(out) :: energy
   real(wp), intent(out) :: gradient(:, :)
   integer, intent(out) :: stat
   logical :: spec_calc

   integer(c_int32_t), allocatable :: cnum(:)
   real(c_double), allocatable :: cxyz(:, :), cgrad(:, :), cqat(:)
   real(c_double) :: cenergy
   integer(c_int32_t) :: cstat, rc
   integer :: nat, unit

   nat = size(num)
   allocate(cnum(nat), cxyz(3, nat), cgrad(3, nat), cqat(nat))
   cnum(:) = int(num(:), c_int32_t)
   cxyz(:, :) = real(xyz(:, 1:nat), c_double)
   cgrad(:, :) = 0.0_c_double
   cqat(:) = 0.0_c_double
   cenergy = 0.0_c_double
   cstat = 0

   rc = c_get_xtb_egrad(int(nat, c_int32_t), cnum, cxyz, int(charge, c_int32_t), &
      & int(multiplicity, c_int32_t), int(method%id, c_int32_t), real(etemp, c_double), &
      & cqat, cenergy, cgrad, cstat)

   stat = int(cstat)
   if (rc /= 0 .and. stat == 0) stat = int(rc)
   if (stat /= 0 .and. stat /= 5) stat = stat_fatal

   energy = real(cenergy, wp)
   gradient(:, 1:nat) = real(cgrad, wp)
   qat(1:nat) = real(cqat, wp)

   open(newunit=unit, file=output_file)
   select case(stat)
   case(0)
      write(unit, '(a)') "[Info] GFN2-xTB calculation on the GPU (libqcxms_gpu)"
   case(5)
      write(unit, '(a)') "[Fatal] Unknown method requested for calculation on the GPU"
   case default
      write(unit, '(a)') "[Fatal] Singlepoint calculation failed"
      write(unit, '(a,a)') "-> ", last_error()
   end select
   if (spec_calc .and. stat == 0) then
      if (.not. write_mo_spectrum(nat)) then
         write(unit, '(a)') "[Fatal] MO spectrum failed"
         write(unit, '(a,a)') "-> ", last_error()
         stat = stat_fatal
      end if
   end if
   close(unit)
end subroutine get_xtb_egrad

!> tmp.mspec and qcxms.Mspec.tbxtb of the last structure evaluated, in the
!> layout src/mo_energ.f90:54-76 writes (read back by readpopmo,
!> src/mo_energ.f90:82-107): orbital energies in eV, occupations, and the
!> populations qmo(k, atom) normalised per orbital.
logical function write_mo_spectrum(nat) result(ok)
   integer, intent(in) :: nat
   integer(c_int32_t) :: cap, nao, ihomo
   real(c_double), allocatable :: emo(:), focc(:), qmo(:, :)
   integer :: io_tmp, io_mspec, j, k

   cap = int(9 * nat, c_int32_t)
   allocate(emo(cap), focc(cap), qmo(nat, cap))
   ok = c_get_xtb_mo_spectrum(cap, nao, ihomo, emo, focc, qmo) == 0
   if (.not. ok) return

   open(newunit=io_tmp, file='tmp.mspec', status='replace')
   open(newunit=io_mspec, file='qcxms.Mspec.tbxtb', status='replace')
   write(io_mspec, *) nao, ihomo
   do k = 1, nao
      write(io_mspec, *)
      write(io_mspec, '(1(i3, 1x, f10.3))') k, emo(k) * autoev
      write(io_mspec, '(1(1x, f6.2))') focc(k)
      write(io_mspec, '(10(1x, f6.2))') (qmo(j, k) * 100.0_wp, j = 1, nat)
   end do
   write(io_tmp, *) nao, ihomo
   do k = 1, nao
      write(io_tmp, *) emo(k) * autoev
      write(io_tmp, *) focc(k)
      do j = 1, nat
         write(io_tmp, *) qmo(j, k)
      end do
   end do
   close(io_tmp)
   close(io_mspec)
end function write_mo_spectrum

!> Message of the library's last failure on this thread
function last_error() result(msg)
   character(len=:), allocatable :: msg
   type(c_ptr) :: p
   character(kind=c_char), pointer :: chars(:)
   integer :: n
   p = c_last_error()
   msg = ""
   if (.not. c_associated(p)) return
   call c_f_pointer(p, chars, [4096])
   n = 0
   do while (n < 4096)
      if (chars(n + 1) == achar(0)) exit
      n = n + 1
   end do
   allocate(character(len=n) :: msg)
   msg = transfer(chars(1:n), msg)
end function last_error

end module qcxms_tblite
