!> Small Fortran caller of the drop-in module, used by the tests: it calls
!> get_xtb_egrad exactly as src/iniqm.f90:644-645 (spec_calc false) and
!> src/iniqm.f90:409-410 (getspec, spec_calc true) do, once per record.
!> stdin, one or more records:
!>         nat charge multiplicity method etemp spec_calc(0/1)
!>         nat lines "Z x y z" (Bohr)
!> stdout, per record: stat, energy, then nat lines "qat gx gy gz"
program egrad_driver
   use qcxms_tblite, only : get_xtb_egrad, gfn1_xtb, gfn2_xtb, ipea1_xtb, method_selector
   implicit none
   integer, parameter :: wp = selected_real_kind(15)
   integer :: nat, chrg, mult, mid, ispec, stat, i, ios
   real(wp) :: etemp, energy
   integer, allocatable :: iat(:)
   real(wp), allocatable :: xyz(:, :), grad(:, :), qat(:)
   type(method_selector) :: method
   logical :: spec_calc

   do
      read(*, *, iostat=ios) nat, chrg, mult, mid, etemp, ispec
      if (ios /= 0) exit
      if (allocated(iat)) deallocate(iat, xyz, grad, qat)
      allocate(iat(nat), xyz(3, nat), grad(3, nat), qat(nat))
      do i = 1, nat
         read(*, *) iat(i), xyz(1, i), xyz(2, i), xyz(3, i)
      end do
      select case(mid)
      case(1)
         method = gfn1_xtb
      case(11)
         method = ipea1_xtb
      case default
         method = gfn2_xtb
      end select
      spec_calc = ispec /= 0
      call get_xtb_egrad(iat, xyz, chrg, mult, method, etemp, "xtb.last", qat, energy, grad, &
         & stat, spec_calc)
      write(*, '(i0)') stat
      write(*, '(es25.17)') energy
      do i = 1, nat
         write(*, '(4es25.17)') qat(i), grad(1, i), grad(2, i), grad(3, i)
      end do
   end do
end program egrad_driver
