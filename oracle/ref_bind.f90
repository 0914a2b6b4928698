! C-callable wrappers around QCxMS's own Fortran routines, compiled from the
! reference sources in place (see oracle/Makefile) into oracle/_ref/.
! Test infrastructure only: the test suite uses these to pin the C oracle and
! to write golden fixtures.  Nothing in the product links this.
!
! Wrapped routines (reference file:line):
!   fragment_structure   src/fragments.f90:93-182
!   fragmass             src/fragments.f90:10-84
!   ekinet               src/mdinit.f90:57-76
!   impactscale          src/impact.f90:12-55
!   Rad                  src/covalent_radii.f90:89-114
!   covalent_radius_d3   src/covalent_radii.f90:83-84
!   atomic_mass          src/atomic_masses.f90:58-67
!   zeta/dzeta/trapzd/cngw/d4dim/d4/covncoord/dispgrad
!                        src/dftd4.f90:292-1337, 1463-1780
!   gam, gam3            src/dftd4.f90:38-72 (the tables eself uses)
!   ref_res_record       a record written with the format of
!                        src/write_fragments.f90:402-441 (the qcxms.res line)
!   ref_write_start      qcxms.start with the write statements of
!                        src/utility.f90:358-369 (wrstart)
!   ref_write_gs         qcxms.gs with the write statements of
!                        src/md.f90:151-152, 383-386
!   getmaxiee, getieeab, gauss0, poiss0
!                        src/iee.f90:12-135
!   calctrelax           src/impact.f90:58-81
module qcxms_ref_bind
   use iso_c_binding
   use xtb_mctc_accuracy, only: wp
   use qcxms_fragments, only: fragment_structure, fragmass
   use qcxms_mdinit, only: ekinet
   use qcxms_impact, only: impactscale, calctrelax
   use qcxms_iee, only: getmaxiee, getieeab, gauss0, poiss0
   use covalent_radii, only: Rad, covalent_radius_d3
   use xtb_atomic_masses, only: atomic_mass
   use dftd4
   implicit none
contains

subroutine ref_fragment_structure(nat, oz, xyz, rcut, at1, at2, frag) bind(C)
   integer(c_int), value :: nat, at1, at2
   integer(c_int), intent(in) :: oz(nat)
   real(c_double), intent(in) :: xyz(3, nat)
   real(c_double), value :: rcut
   integer(c_int), intent(out) :: frag(nat)
   integer :: oz_(nat), frag_(nat)
   real(wp) :: r
   oz_ = oz
   r = rcut
   call fragment_structure(int(nat), oz_, xyz, r, int(at1), int(at2), frag_)
   frag = frag_
end subroutine

subroutine ref_fragmass(nat, iat, list, mass, imass, nfrag, fragm, fragat) bind(C)
   integer(c_int), value :: nat
   integer(c_int), intent(in) :: iat(nat), list(nat), imass(nat)
   real(c_double), intent(in) :: mass(nat)
   integer(c_int), intent(out) :: nfrag
   real(c_double), intent(out) :: fragm(10)
   integer(c_int), intent(out) :: fragat(200, 10)
   integer :: iat_(nat), list_(nat), imass_(nat), nfrag_, fragat_(200, 10)
   real(wp) :: mass_(nat), fragm_(10)
   character(len=80) :: fragf(max(nat, 10))
   iat_ = iat; list_ = list; imass_ = imass; mass_ = mass
   fragm_ = 0.0_wp
   fragat_ = 0
   call fragmass(int(nat), iat_, list_, mass_, imass_, nfrag_, fragm_, fragf, fragat_)
   nfrag = nfrag_
   fragm = fragm_
   fragat = fragat_
end subroutine

subroutine ref_ekinet(nat, velo, mass, e, t) bind(C)
   integer(c_int), value :: nat
   real(c_double), intent(in) :: velo(3, nat), mass(nat)
   real(c_double), intent(out) :: e, t
   real(wp) :: e_, t_
   call ekinet(int(nat), velo, mass, e_, t_)
   e = e_
   t = t_
end subroutine

subroutine ref_impactscale(nuc, velo, mass, velof, eimp, ff, e0) bind(C)
   integer(c_int), value :: nuc
   real(c_double), intent(inout) :: velo(3, nuc)
   real(c_double), intent(in) :: mass(nuc), velof(nuc)
   real(c_double), value :: eimp, ff, e0
   real(wp) :: v_(3, nuc), m_(nuc), vf_(nuc), eimp_, ff_, e0_
   v_ = velo; m_ = mass; vf_ = velof; eimp_ = eimp; ff_ = ff; e0_ = e0
   call impactscale(int(nuc), v_, m_, vf_, eimp_, ff_, e0_)
   velo = v_
end subroutine

function ref_rad(z) result(r) bind(C)
   integer(c_int), value :: z
   real(c_double) :: r
   r = Rad(z)
end function

function ref_covrad_d3(z) result(r) bind(C)
   integer(c_int), value :: z
   real(c_double) :: r
   r = covalent_radius_d3(z)
end function

function ref_atomic_mass(z) result(m) bind(C)
   integer(c_int), value :: z
   real(c_double) :: m
   m = atomic_mass(int(z))
end function

function ref_zeta(a, c, qref, qmod) result(z) bind(C)
   real(c_double), value :: a, c, qref, qmod
   real(c_double) :: z
   z = zeta(a, c, qref, qmod)
end function

function ref_dzeta(a, c, qref, qmod) result(z) bind(C)
   real(c_double), value :: a, c, qref, qmod
   real(c_double) :: z
   z = dzeta(a, c, qref, qmod)
end function

function ref_trapzd(pol) result(t) bind(C)
   real(c_double), intent(in) :: pol(23)
   real(c_double) :: t
   t = trapzd(pol)
end function

function ref_cngw(wf, cn, cnref) result(w) bind(C)
   real(c_double), value :: wf, cn, cnref
   real(c_double) :: w
   w = cngw(wf, cn, cnref)
end function

!> Reference-system table as d4dim leaves it (mode selects the charge set).
subroutine ref_d4dim(nat, at, g_a, g_c, mode, ndim) bind(C)
   integer(c_int), value :: nat, mode
   integer(c_int), intent(in) :: at(nat)
   real(c_double), value :: g_a, g_c
   integer(c_int), intent(out) :: ndim
   integer :: at_(nat), nd
   at_ = at
   call d4dim(int(nat), at_, g_a, g_c, int(mode), nd)
   ndim = nd
end subroutine

subroutine ref_d4_tables(z, nref, refq_out, refcovcn_out, refc_out, refal_out) bind(C)
   integer(c_int), value :: z
   integer(c_int), intent(out) :: nref
   real(c_double), intent(out) :: refq_out(7), refcovcn_out(7), refal_out(23, 7)
   integer(c_int), intent(out) :: refc_out(7)
   nref = refn(z)
   refq_out = refq(:, z)
   refcovcn_out = refcovcn(:, z)
   refc_out = refc(:, z)
   refal_out = refal(:, :, z)
end subroutine

subroutine ref_d4(nat, ndim, at, wf, g_a, g_c, covcn, gw, c6abns) bind(C)
   integer(c_int), value :: nat, ndim
   integer(c_int), intent(in) :: at(nat)
   real(c_double), value :: wf, g_a, g_c
   real(c_double), intent(in) :: covcn(nat)
   real(c_double), intent(out) :: gw(ndim), c6abns(ndim, ndim)
   integer :: at_(nat)
   at_ = at
   call d4(int(nat), int(ndim), at_, wf, g_a, g_c, covcn, gw, c6abns)
end subroutine

subroutine ref_covncoord(nat, at, xyz, cn, thr) bind(C)
   integer(c_int), value :: nat
   integer(c_int), intent(in) :: at(nat)
   real(c_double), intent(in) :: xyz(3, nat)
   real(c_double), value :: thr
   real(c_double), intent(out) :: cn(nat)
   integer :: at_(nat)
   at_ = at
   call covncoord(int(nat), at_, xyz, cn, thr)
end subroutine

!> Full D4 energy (two-body at charges q + approximate ATM at q=0) and gradient.
subroutine ref_dispgrad(nat, ndim, at, q, xyz, s6, s8, s9, a1, a2, alp, wf, g_a, g_c, &
      & g, eout) bind(C)
   integer(c_int), value :: nat, ndim, alp
   integer(c_int), intent(in) :: at(nat)
   real(c_double), intent(in) :: q(nat), xyz(3, nat)
   real(c_double), value :: s6, s8, s9, a1, a2, wf, g_a, g_c
   real(c_double), intent(inout) :: g(3, nat)
   real(c_double), intent(out) :: eout
   integer :: at_(nat)
   real(wp), allocatable :: gw(:), c6abns(:, :), cn(:)
   at_ = at
   allocate(gw(ndim), c6abns(ndim, ndim), cn(nat))
   call covncoord(int(nat), at_, xyz, cn, 1600.0_wp)
   call d4(int(nat), int(ndim), at_, wf, g_a, g_c, cn, gw, c6abns)
   call dispgrad(int(nat), int(ndim), at_, q, xyz, s6, s8, s9, a1, a2, int(alp), wf, &
      & g_a, g_c, c6abns, 3, g, eout)
end subroutine

function ref_gam(z) result(g) bind(C)
   integer(c_int), value :: z
   real(c_double) :: g
   g = gam(z)
end function

function ref_gam3(z) result(g) bind(C)
   integer(c_int), value :: z
   real(c_double) :: g
   g = gam3(z)
end function

! One qcxms.res record, items as write_fragments.f90 passes them: EI
! (icoll < 0) without the collision number, CID with it.  out receives the
! record blank-padded to outlen characters.
subroutine ref_res_record(q, mchrg, itrj, icoll, isec, j, l, idum2, idum1, out, outlen) bind(C)
   real(c_double), value :: q
   integer(c_int), value :: mchrg, itrj, icoll, isec, j, l, outlen
   integer(c_int), intent(in) :: idum2(*), idum1(*)
   character(kind=c_char), intent(out) :: out(outlen)
   character(len=400) :: line
   integer :: m
   line = ''
   if (icoll < 0) then
      write(line,'(F10.7,i3,2i5,2i2,2x,i3,2x,20(i4,i3))') q, mchrg, itrj, isec, j, &
         & l, (idum2(m), idum1(m), m = 1, l)
   else
      write(line,'(F10.7,i3,2i5,2i2,2x,i3,2x,20(i4,i3))') q, mchrg, itrj, icoll, isec, &
         & j, l, (idum2(m), idum1(m), m = 1, l)
   endif
   do m = 1, outlen
      out(m) = line(m:m)
   enddo
end subroutine

subroutine ref_write_start(path, plen, traj, nat, velo, velof, eimp, tadd) bind(C)
   integer(c_int), value :: plen, traj, nat
   character(kind=c_char), intent(in) :: path(plen)
   real(c_double), intent(in) :: velo(3, nat), velof(nat)
   real(c_double), value :: eimp, tadd
   character(len=plen) :: fname
   integer :: u, j
   do j = 1, plen
      fname(j:j) = path(j)
   enddo
   open(file=fname, newunit=u)
   write(u,'(i4)') traj
   write(u,'(D22.14)') eimp
   write(u,'(D22.14)') tadd
   do j = 1, nat
      write(u,'(7D22.14)') velo(1:3, j), velof(j)
   enddo
   close(u)
end subroutine

subroutine ref_write_gs(path, plen, nmax, ndump, nat, xyz, velo) bind(C)
   integer(c_int), value :: plen, nmax, ndump, nat
   character(kind=c_char), intent(in) :: path(plen)
   real(c_double), intent(in) :: xyz(3, nat, ndump), velo(3, nat, ndump)
   character(len=plen) :: fname
   integer :: u, i, j, k
   do j = 1, plen
      fname(j:j) = path(j)
   enddo
   open(file=fname, newunit=u)
   write(u,*) nmax
   do k = 1, ndump
      do i = 1, nat
         write(u,'(6d16.8)') (xyz(j, i, k), j = 1, 3), (velo(j, i, k), j = 1, 3)
      enddo
   enddo
   close(u)
end subroutine

subroutine ref_getmaxiee(a, b, ieeel, ityp, exc, ieemax, pmax, e_avg) bind(C)
   real(c_double), value :: a, b, ieeel, exc
   integer(c_int), value :: ityp
   real(c_double), intent(out) :: ieemax, pmax, e_avg
   real(wp) :: a_, b_, el_, ex_
   a_ = a; b_ = b; el_ = ieeel; ex_ = exc
   call getmaxiee(a_, b_, el_, int(ityp), ex_, ieemax, pmax, e_avg)
end subroutine

subroutine ref_getieeab(ieeel, ityp, exc, nbnd, ieeatm, a, b) bind(C)
   real(c_double), value :: ieeel, exc, ieeatm
   integer(c_int), value :: ityp, nbnd
   real(c_double), intent(out) :: a, b
   real(wp) :: el_, ex_, at_
   el_ = ieeel; ex_ = exc; at_ = ieeatm
   call getieeab(a, b, el_, int(ityp), ex_, int(nbnd), at_)
end subroutine

function ref_gauss0(a, b, ieeel, x) result(p) bind(C)
   real(c_double), value :: a, b, ieeel, x
   real(c_double) :: p
   real(wp) :: a_, b_, el_, x_
   a_ = a; b_ = b; el_ = ieeel; x_ = x
   call gauss0(a_, b_, el_, x_, p)
end function

function ref_poiss0(a, b, ieeel, x) result(p) bind(C)
   real(c_double), value :: a, b, ieeel, x
   real(c_double) :: p
   real(wp) :: a_, b_, el_, x_
   a_ = a; b_ = b; el_ = ieeel; x_ = x
   call poiss0(a_, b_, el_, x_, p)
end function

function ref_calctrelax(n, emo, na, i, trelax) result(t) bind(C)
   integer(c_int), value :: n, na, i
   real(c_double), intent(in) :: emo(10 * n)
   real(c_double), value :: trelax
   real(c_double) :: t
   real(wp) :: e_(10 * n), tr_
   e_ = emo
   tr_ = trelax
   call calctrelax(int(n), e_, int(na), int(i), tr_, t)
end function

end module qcxms_ref_bind
