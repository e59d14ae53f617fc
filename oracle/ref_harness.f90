! ORACLE-SIDE TEST INFRASTRUCTURE -- not part of the product.
!
! C-callable harness around the reference Fortran (brichjnh/hfcode), compiled
! from the unmodified reference sources by oracle/Makefile into
! oracle/_ref/libhfref.so.  Only tests/, the smoke check and bench.py's CPU
! baseline leg load it, and only as the checker / the timed CPU reference.
!
! It drives the reference set-up exactly as hf.f90:24-33 does
! (read_infile -> nuclear_repulsion -> read_basfile -> spread_basis ->
! precalcs_kabs -> cmpt1e) and then exposes:
!   * the Fock-build routines of the hot path: fock_build_initial_diag
!     (fock_build_initial_diag.f90:1), fock_build_diag (fock_build_diag.f90:1),
!     fock_build_non_diag (fock_build_non_diag.f90:1);
!   * the per-class ERI routines i2e_* (s_sp_l_terms.f90, s_p_d_terms.f90,
!     s_d_l_terms.f90) and the Boys function boys_n (boys.f90:115);
!   * evaluate_maxdm (evaluate_maxdm.f90:1);
!   * the module state the drop-in boundary marshals (molprops.f90:37-168):
!     shells, shell-pair lists, Schwarz tables, one-electron matrices.
module hfref_harness
use iso_c_binding
use nrtype
use molprops
use boys
use s_sp_l_terms
use s_p_d_terms
use s_d_l_terms
implicit none

contains

subroutine c2f(cstr, fstr)
  character(kind=c_char), dimension(*), intent(in) :: cstr
  character(len=*), intent(out) :: fstr
  integer :: i
  fstr = ' '
  do i = 1, len(fstr)
    if (cstr(i) == c_null_char) exit
    fstr(i:i) = cstr(i)
  end do
end subroutine c2f

! Set-up of one molecule, in the order of hf.f90:24-33.  File names are
! relative to the caller's working directory and at most 30 characters
! (read_infile.f90:4, hf.f90:6).  Returns nb.
integer(c_int) function hfref_load(cin, cbas, clog) bind(C, name="hfref_load")
  character(kind=c_char), dimension(*), intent(in) :: cin, cbas, clog
  character(len=30) :: infile, basfile
  character(len=256) :: logfile
  call c2f(cin, infile)
  call c2f(cbas, basfile)
  call c2f(clog, logfile)
  open(unit=9, file=trim(logfile))
  call read_infile(infile)
  call nuclear_repulsion()
  call read_basfile(basfile)
  call spread_basis()
  call precalcs_kabs()
  call cmpt1e()
  flush(9)
  hfref_load = nb
end function hfref_load

subroutine hfref_dims(d) bind(C, name="hfref_dims")
  integer(c_int), intent(out) :: d(32)
  d = 0
  d(1) = natom ; d(2) = nb ; d(3) = nsh ; d(4) = nocc ; d(5) = nab
  d(6) = nshs ; d(7) = nshp ; d(8) = nshd ; d(9) = nshl ; d(10) = maxprim
  d(11) = nss ; d(12) = nsp ; d(13) = nsd ; d(14) = nsl ; d(15) = npp
  d(16) = npd ; d(17) = npl ; d(18) = ndd ; d(19) = ndl ; d(20) = nll
end subroutine hfref_dims

real(c_double) function hfref_unuc() bind(C, name="hfref_unuc")
  hfref_unuc = unuc
end function hfref_unuc

! Shell table: type code (1=s,2=p,3=d,12=l), primitive count, atom, first AO
! (1-based), centre (bohr), exponents, contraction coefficients d_jk and the
! L-shell p/s coefficient ratios djkl (zero for non-L shells).
subroutine hfref_shells(ityp, inz, iat, iao, xyz, zet, cof, cofl) bind(C, name="hfref_shells")
  integer(c_int), intent(out) :: ityp(nsh), inz(nsh), iat(nsh), iao(nsh)
  real(c_double), intent(out) :: xyz(3,nsh), zet(maxprim,nsh), cof(maxprim,nsh), cofl(maxprim,nsh)
  integer :: i
  ityp = shtyp ; inz = nzet ; iat = atsh ; iao = aosh
  xyz = xsh
  zet = 0._dp ; cof = 0._dp ; cofl = 0._dp
  do i = 1, nsh
    zet(1:nzet(i),i) = zeta(1:nzet(i),i)
    cof(1:nzet(i),i) = djk(1:nzet(i),i)
    if (shtyp(i) == 12) cofl(1:nzet(i),i) = djkl(1:nzet(i),i)
  end do
end subroutine hfref_shells

! Pair list, Schwarz values, Schwarz order and truncation index of pair type
! t = 1..10 (ss, sp, sd, sl, pp, pd, pl, dd, dl, ll).  Valid after
! fock_build_initial_diag has run (fock_build_initial_diag.f90:51-52).
integer(c_int) function hfref_pairs(t, prs, sch, ord, last) bind(C, name="hfref_pairs")
  integer(c_int), value :: t
  integer(c_int), intent(out) :: prs(2,*), ord(*), last
  real(c_double), intent(out) :: sch(*)
  integer :: n
  n = 0 ; last = 0
  select case (t)
  case (1)
    n = nss ; if (n > 0) then ; prs(:,1:n) = ssprs ; sch(1:n) = schwss ; ord(1:n) = orderss ; last = lastss ; end if
  case (2)
    n = nsp ; if (n > 0) then ; prs(:,1:n) = spprs ; sch(1:n) = schwsp ; ord(1:n) = ordersp ; last = lastsp ; end if
  case (3)
    n = nsd ; if (n > 0) then ; prs(:,1:n) = sdprs ; sch(1:n) = schwsd ; ord(1:n) = ordersd ; last = lastsd ; end if
  case (4)
    n = nsl ; if (n > 0) then ; prs(:,1:n) = slprs ; sch(1:n) = schwsl ; ord(1:n) = ordersl ; last = lastsl ; end if
  case (5)
    n = npp ; if (n > 0) then ; prs(:,1:n) = ppprs ; sch(1:n) = schwpp ; ord(1:n) = orderpp ; last = lastpp ; end if
  case (6)
    n = npd ; if (n > 0) then ; prs(:,1:n) = pdprs ; sch(1:n) = schwpd ; ord(1:n) = orderpd ; last = lastpd ; end if
  case (7)
    n = npl ; if (n > 0) then ; prs(:,1:n) = plprs ; sch(1:n) = schwpl ; ord(1:n) = orderpl ; last = lastpl ; end if
  case (8)
    n = ndd ; if (n > 0) then ; prs(:,1:n) = ddprs ; sch(1:n) = schwdd ; ord(1:n) = orderdd ; last = lastdd ; end if
  case (9)
    n = ndl ; if (n > 0) then ; prs(:,1:n) = dlprs ; sch(1:n) = schwdl ; ord(1:n) = orderdl ; last = lastdl ; end if
  case (10)
    n = nll ; if (n > 0) then ; prs(:,1:n) = llprs ; sch(1:n) = schwll ; ord(1:n) = orderll ; last = lastll ; end if
  end select
  hfref_pairs = n
end function hfref_pairs

subroutine hfref_onee(s, t, v) bind(C, name="hfref_onee")
  real(c_double), intent(out) :: s(nb,nb), t(nb,nb), v(nb,nb)
  s = s1e ; t = t1e ; v = v1e
end subroutine hfref_onee

subroutine hfref_initial_diag(f, p, thr) bind(C, name="hfref_initial_diag")
  real(c_double), intent(inout) :: f(nb,nb)
  real(c_double), intent(in) :: p(nb,nb)
  real(c_double), value :: thr
  call fock_build_initial_diag(f, p, thr)
end subroutine hfref_initial_diag

subroutine hfref_diag(f, p, thr) bind(C, name="hfref_diag")
  real(c_double), intent(inout) :: f(nb,nb)
  real(c_double), intent(in) :: p(nb,nb)
  real(c_double), value :: thr
  call fock_build_diag(f, p, thr)
end subroutine hfref_diag

subroutine hfref_non_diag(f, p, thr) bind(C, name="hfref_non_diag")
  real(c_double), intent(inout) :: f(nb,nb)
  real(c_double), intent(in) :: p(nb,nb)
  real(c_double), value :: thr
  call fock_build_non_diag(f, p, thr)
end subroutine hfref_non_diag

subroutine hfref_maxdm(p, maxdm) bind(C, name="hfref_maxdm")
  real(c_double), intent(in) :: p(nb,nb)
  real(c_double), intent(out) :: maxdm(nab)
  call evaluate_maxdm(p, maxdm)
end subroutine hfref_maxdm

subroutine hfref_boys(n, x, fn) bind(C, name="hfref_boys")
  integer(c_int), value :: n
  real(c_double), value :: x
  real(c_double), intent(out) :: fn(0:n)
  real(dp) :: f1(0:1), f2(0:2)
  if (n == 0) then
    call boys_0(x, fn(0))
  else if (n == 1) then
    call boys_1(x, f1) ; fn(0:1) = f1
  else if (n == 2) then
    call boys_2(x, f2) ; fn(0:2) = f2
  else
    call boys_n(n, x, fn)
  end if
end subroutine hfref_boys

! One contracted block from the reference class routine `cls` (0..45, in the
! order of the class loops of fock_build_non_diag.f90), shells 1-based.  The
! block is written column-major with the dimensions of the routine's dummy
! argument (s shells contribute no dimension).
subroutine hfref_eri(cls, sh1, sh2, sh3, sh4, buf) bind(C, name="hfref_eri")
  integer(c_int), value :: cls, sh1, sh2, sh3, sh4
  real(c_double), intent(out) :: buf(625)
  buf = 0._dp
  select case (cls)
  case (0)
    call i2e_ssss(sh1,sh2,sh3,sh4,buf(1))   ! block scalar, 1 values
  case (1)
    call i2e_sssp(sh1,sh2,sh3,sh4,buf)   ! block 3, 3 values
  case (2)
    call i2e_sssd(sh1,sh2,sh3,sh4,buf)   ! block 5, 5 values
  case (3)
    call i2e_sssl(sh1,sh2,sh3,sh4,buf)   ! block 4, 4 values
  case (4)
    call i2e_sspp(sh1,sh2,sh3,sh4,buf)   ! block 3x3, 9 values
  case (5)
    call i2e_sspd(sh1,sh2,sh3,sh4,buf)   ! block 3x5, 15 values
  case (6)
    call i2e_sspl(sh1,sh2,sh3,sh4,buf)   ! block 3x4, 12 values
  case (7)
    call i2e_ssdd(sh1,sh2,sh3,sh4,buf)   ! block 5x5, 25 values
  case (8)
    call i2e_ssdl(sh1,sh2,sh3,sh4,buf)   ! block 5x4, 20 values
  case (9)
    call i2e_ssll(sh1,sh2,sh3,sh4,buf)   ! block 4x4, 16 values
  case (10)
    call i2e_spsp(sh1,sh2,sh3,sh4,buf)   ! block 3x3, 9 values
  case (11)
    call i2e_spsd(sh1,sh2,sh3,sh4,buf)   ! block 3x5, 15 values
  case (12)
    call i2e_spsl(sh1,sh2,sh3,sh4,buf)   ! block 3x4, 12 values
  case (13)
    call i2e_sppp(sh1,sh2,sh3,sh4,buf)   ! block 3x3x3, 27 values
  case (14)
    call i2e_sppd(sh1,sh2,sh3,sh4,buf)   ! block 3x3x5, 45 values
  case (15)
    call i2e_sppl(sh1,sh2,sh3,sh4,buf)   ! block 3x3x4, 36 values
  case (16)
    call i2e_spdd(sh1,sh2,sh3,sh4,buf)   ! block 3x5x5, 75 values
  case (17)
    call i2e_spll(sh1,sh2,sh3,sh4,buf)   ! block 3x4x4, 48 values
  case (18)
    call i2e_sdsd(sh1,sh2,sh3,sh4,buf)   ! block 5x5, 25 values
  case (19)
    call i2e_sdsl(sh1,sh2,sh3,sh4,buf)   ! block 5x4, 20 values
  case (20)
    call i2e_sdpp(sh1,sh2,sh3,sh4,buf)   ! block 5x3x3, 45 values
  case (21)
    call i2e_sdpd(sh1,sh2,sh3,sh4,buf)   ! block 5x3x5, 75 values
  case (22)
    call i2e_sddd(sh1,sh2,sh3,sh4,buf)   ! block 5x5x5, 125 values
  case (23)
    call i2e_sddl(sh1,sh2,sh3,sh4,buf)   ! block 5x5x4, 100 values
  case (24)
    call i2e_sdll(sh1,sh2,sh3,sh4,buf)   ! block 5x4x4, 80 values
  case (25)
    call i2e_slsl(sh1,sh2,sh3,sh4,buf)   ! block 4x4, 16 values
  case (26)
    call i2e_slpp(sh1,sh2,sh3,sh4,buf)   ! block 4x3x3, 36 values
  case (27)
    call i2e_slpl(sh1,sh2,sh3,sh4,buf)   ! block 4x3x4, 48 values
  case (28)
    call i2e_sldd(sh1,sh2,sh3,sh4,buf)   ! block 4x5x5, 100 values
  case (29)
    call i2e_sldl(sh1,sh2,sh3,sh4,buf)   ! block 4x5x4, 80 values
  case (30)
    call i2e_slll(sh1,sh2,sh3,sh4,buf)   ! block 4x4x4, 64 values
  case (31)
    call i2e_pppp(sh1,sh2,sh3,sh4,buf)   ! block 3x3x3x3, 81 values
  case (32)
    call i2e_pppd(sh1,sh2,sh3,sh4,buf)   ! block 3x3x3x5, 135 values
  case (33)
    call i2e_pppl(sh1,sh2,sh3,sh4,buf)   ! block 3x3x3x4, 108 values
  case (34)
    call i2e_ppdd(sh1,sh2,sh3,sh4,buf)   ! block 3x3x5x5, 225 values
  case (35)
    call i2e_ppll(sh1,sh2,sh3,sh4,buf)   ! block 3x3x4x4, 144 values
  case (36)
    call i2e_pdpd(sh1,sh2,sh3,sh4,buf)   ! block 3x5x3x5, 225 values
  case (37)
    call i2e_pddd(sh1,sh2,sh3,sh4,buf)   ! block 3x5x5x5, 375 values
  case (38)
    call i2e_plpl(sh1,sh2,sh3,sh4,buf)   ! block 3x4x3x4, 144 values
  case (39)
    call i2e_plll(sh1,sh2,sh3,sh4,buf)   ! block 3x4x4x4, 192 values
  case (40)
    call i2e_dddd(sh1,sh2,sh3,sh4,buf)   ! block 5x5x5x5, 625 values
  case (41)
    call i2e_dddl(sh1,sh2,sh3,sh4,buf)   ! block 5x5x5x4, 500 values
  case (42)
    call i2e_ddll(sh1,sh2,sh3,sh4,buf)   ! block 5x5x4x4, 400 values
  case (43)
    call i2e_dldl(sh1,sh2,sh3,sh4,buf)   ! block 5x4x5x4, 400 values
  case (44)
    call i2e_dlll(sh1,sh2,sh3,sh4,buf)   ! block 5x4x4x4, 320 values
  case (45)
    call i2e_llll(sh1,sh2,sh3,sh4,buf)   ! block 4x4x4x4, 256 values
  end select
end subroutine hfref_eri

! Release every allocatable of module molprops so that another molecule can be
! loaded in the same process.
subroutine hfref_free() bind(C, name="hfref_free")
  close(9)
  if (allocated(prsh)) deallocate(prsh)
  if (allocated(ind2)) deallocate(ind2)
  if (allocated(ioff)) deallocate(ioff)
  if (allocated(revind2)) deallocate(revind2)
  if (allocated(atlabels)) deallocate(atlabels)
  if (allocated(cij)) deallocate(cij)
  if (allocated(moen)) deallocate(moen)
  if (allocated(atcoords)) deallocate(atcoords)
  if (allocated(angcoords)) deallocate(angcoords)
  if (allocated(nuccharges)) deallocate(nuccharges)
  if (allocated(atznumber)) deallocate(atznumber)
  if (allocated(basels)) deallocate(basels)
  if (allocated(bastypes)) deallocate(bastypes)
  if (allocated(basnshell)) deallocate(basnshell)
  if (allocated(basnspd)) deallocate(basnspd)
  if (allocated(basnzet)) deallocate(basnzet)
  if (allocated(baszeta)) deallocate(baszeta)
  if (allocated(basdjk)) deallocate(basdjk)
  if (allocated(basdjkl)) deallocate(basdjkl)
  if (allocated(basdmat)) deallocate(basdmat)
  if (allocated(shtyp)) deallocate(shtyp)
  if (allocated(nzet)) deallocate(nzet)
  if (allocated(atsh)) deallocate(atsh)
  if (allocated(nshat)) deallocate(nshat)
  if (allocated(nshtat)) deallocate(nshtat)
  if (allocated(shat)) deallocate(shat)
  if (allocated(nbfsat)) deallocate(nbfsat)
  if (allocated(bsref)) deallocate(bsref)
  if (allocated(shtyat)) deallocate(shtyat)
  if (allocated(aosh)) deallocate(aosh)
  if (allocated(ssh)) deallocate(ssh)
  if (allocated(psh)) deallocate(psh)
  if (allocated(dsh)) deallocate(dsh)
  if (allocated(lsh)) deallocate(lsh)
  if (allocated(xsh)) deallocate(xsh)
  if (allocated(zeta)) deallocate(zeta)
  if (allocated(djk)) deallocate(djk)
  if (allocated(djkl)) deallocate(djkl)
  if (allocated(gammrats)) deallocate(gammrats)
  if (allocated(boyszeroarg)) deallocate(boyszeroarg)
  if (allocated(boysfuncvals)) deallocate(boysfuncvals)
  if (allocated(boysfuncargs)) deallocate(boysfuncargs)
  if (allocated(boyslrfac)) deallocate(boyslrfac)
  if (allocated(boysmrfac)) deallocate(boysmrfac)
  if (allocated(atao)) deallocate(atao)
  if (allocated(aohuck)) deallocate(aohuck)
  if (allocated(shao)) deallocate(shao)
  if (allocated(aotyp)) deallocate(aotyp)
  if (allocated(ssprs)) deallocate(ssprs)
  if (allocated(spprs)) deallocate(spprs)
  if (allocated(sdprs)) deallocate(sdprs)
  if (allocated(slprs)) deallocate(slprs)
  if (allocated(ppprs)) deallocate(ppprs)
  if (allocated(pdprs)) deallocate(pdprs)
  if (allocated(plprs)) deallocate(plprs)
  if (allocated(ddprs)) deallocate(ddprs)
  if (allocated(dlprs)) deallocate(dlprs)
  if (allocated(llprs)) deallocate(llprs)
  if (allocated(pab)) deallocate(pab)
  if (allocated(pm1)) deallocate(pm1)
  if (allocated(hpm1)) deallocate(hpm1)
  if (allocated(qab)) deallocate(qab)
  if (allocated(rab2)) deallocate(rab2)
  if (allocated(r2AB)) deallocate(r2AB)
  if (allocated(Px)) deallocate(Px)
  if (allocated(negab)) deallocate(negab)
  if (allocated(ddab)) deallocate(ddab)
  if (allocated(kab)) deallocate(kab)
  if (allocated(schw)) deallocate(schw)
  if (allocated(schwss)) deallocate(schwss)
  if (allocated(schwsp)) deallocate(schwsp)
  if (allocated(schwsd)) deallocate(schwsd)
  if (allocated(schwsl)) deallocate(schwsl)
  if (allocated(schwpp)) deallocate(schwpp)
  if (allocated(schwpd)) deallocate(schwpd)
  if (allocated(schwpl)) deallocate(schwpl)
  if (allocated(schwdd)) deallocate(schwdd)
  if (allocated(schwdl)) deallocate(schwdl)
  if (allocated(schwll)) deallocate(schwll)
  if (allocated(orderss)) deallocate(orderss)
  if (allocated(ordersp)) deallocate(ordersp)
  if (allocated(ordersd)) deallocate(ordersd)
  if (allocated(ordersl)) deallocate(ordersl)
  if (allocated(orderpp)) deallocate(orderpp)
  if (allocated(orderpd)) deallocate(orderpd)
  if (allocated(orderpl)) deallocate(orderpl)
  if (allocated(orderdd)) deallocate(orderdd)
  if (allocated(orderdl)) deallocate(orderdl)
  if (allocated(orderll)) deallocate(orderll)
  if (allocated(s1e)) deallocate(s1e)
  if (allocated(t1e)) deallocate(t1e)
  if (allocated(v1e)) deallocate(v1e)
end subroutine hfref_free

end module hfref_harness
