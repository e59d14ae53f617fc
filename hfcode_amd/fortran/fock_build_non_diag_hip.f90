! Drop-in replacement for the reference routines
!   fock_build_non_diag(f,p,thr)          fock_build_non_diag.f90:1-10
!   fock_build_non_diag_j_only(f,p,thr)   fock_build_non_diag_j.f90:1
! with the same names, arguments and module state (use molprops), running the
! two-electron build on the GPU through the hfx C ABI (include/hfx.h).
!
! Link this object in place of fock_build_non_diag.o of the reference makefile
! (makefile:30) and add  -L<repo>/hfcode_amd -lhfx -Wl,-rpath,<repo>/hfcode_amd.
! The first call (made by direct_hartreefock right after
! fock_build_initial_diag has filled the Schwarz tables,
! direct_hartreefock.f90:101-103) copies shells, pair lists and Schwarz tables
! into a library context that later calls reuse.  Errors stop the program with
! a message, the reference's own convention (read_infile.f90:13-16).
module hfx_fortran
use iso_c_binding
implicit none

integer(c_int), parameter :: HFX_JK = 0, HFX_J_ONLY = 1

type, bind(C) :: hfx_mol_desc
  integer(c_int) :: nb, nsh, maxprim
  type(c_ptr) :: shtyp, nzet, aosh, xsh, zeta, djk, djkl
  integer(c_int) :: npair(10)
  type(c_ptr) :: prs(10), schw(10), order(10)
  integer(c_int) :: last(10)
end type hfx_mol_desc

interface
  integer(c_int) function hfx_ctx_create(desc, ctx) bind(C, name="hfx_ctx_create")
    import :: hfx_mol_desc, c_int, c_ptr
    type(hfx_mol_desc), intent(in) :: desc
    type(c_ptr), intent(out) :: ctx
  end function hfx_ctx_create
  integer(c_int) function hfx_fock_nd(ctx, p, f, nb, thr, mode) bind(C, name="hfx_fock_nd")
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: p(*)
    real(c_double), intent(inout) :: f(*)
    integer(c_int), value :: nb, mode
    real(c_double), value :: thr
  end function hfx_fock_nd
end interface

type(c_ptr), save :: hfx_ctx = c_null_ptr

contains

! Marshal the molprops state (molprops.f90:93-158) into an hfx context.
subroutine hfx_context()
  use nrtype
  use molprops
  type(hfx_mol_desc) :: d
  integer(c_int), allocatable, target :: ityp(:), inz(:), iao(:)
  real(c_double), allocatable, target :: xyz(:,:), zet(:,:), cof(:,:), cofl(:,:)
  integer(c_int), allocatable, target :: p1(:,:), p2(:,:), p3(:,:), p4(:,:), p5(:,:), p6(:,:), p7(:,:), p8(:,:), &
                                         p9(:,:), p10(:,:)
  real(c_double), allocatable, target :: s1(:), s2(:), s3(:), s4(:), s5(:), s6(:), s7(:), s8(:), s9(:), s10(:)
  integer(c_int), allocatable, target :: o1(:), o2(:), o3(:), o4(:), o5(:), o6(:), o7(:), o8(:), o9(:), o10(:)
  integer(c_int) :: rc
  integer :: i

  if (c_associated(hfx_ctx)) return
  allocate(ityp(nsh), inz(nsh), iao(nsh), xyz(3,nsh), zet(maxprim,nsh), cof(maxprim,nsh), cofl(maxprim,nsh))
  ityp = shtyp ; inz = nzet ; iao = aosh ; xyz = xsh
  zet = 0._dp ; cof = 0._dp ; cofl = 0._dp
  do i = 1, nsh
    zet(1:nzet(i),i) = zeta(1:nzet(i),i)
    cof(1:nzet(i),i) = djk(1:nzet(i),i)
    if (shtyp(i) == 12) cofl(1:nzet(i),i) = djkl(1:nzet(i),i)
  end do
  d%nb = nb ; d%nsh = nsh ; d%maxprim = maxprim
  d%shtyp = c_loc(ityp) ; d%nzet = c_loc(inz) ; d%aosh = c_loc(iao)
  d%xsh = c_loc(xyz) ; d%zeta = c_loc(zet) ; d%djk = c_loc(cof) ; d%djkl = c_loc(cofl)
  d%npair = 0 ; d%last = 0
  d%prs = c_null_ptr ; d%schw = c_null_ptr ; d%order = c_null_ptr
  ! pair type order ss sp sd sl pp pd pl dd dl ll (molprops.f90:122-131)
  if (nss > 0) then ; allocate(p1(2,nss), s1(nss), o1(nss)) ; p1 = ssprs ; s1 = schwss ; o1 = orderss
    d%npair(1) = nss ; d%prs(1) = c_loc(p1) ; d%schw(1) = c_loc(s1) ; d%order(1) = c_loc(o1) ; d%last(1) = lastss ; end if
  if (nsp > 0) then ; allocate(p2(2,nsp), s2(nsp), o2(nsp)) ; p2 = spprs ; s2 = schwsp ; o2 = ordersp
    d%npair(2) = nsp ; d%prs(2) = c_loc(p2) ; d%schw(2) = c_loc(s2) ; d%order(2) = c_loc(o2) ; d%last(2) = lastsp ; end if
  if (nsd > 0) then ; allocate(p3(2,nsd), s3(nsd), o3(nsd)) ; p3 = sdprs ; s3 = schwsd ; o3 = ordersd
    d%npair(3) = nsd ; d%prs(3) = c_loc(p3) ; d%schw(3) = c_loc(s3) ; d%order(3) = c_loc(o3) ; d%last(3) = lastsd ; end if
  if (nsl > 0) then ; allocate(p4(2,nsl), s4(nsl), o4(nsl)) ; p4 = slprs ; s4 = schwsl ; o4 = ordersl
    d%npair(4) = nsl ; d%prs(4) = c_loc(p4) ; d%schw(4) = c_loc(s4) ; d%order(4) = c_loc(o4) ; d%last(4) = lastsl ; end if
  if (npp > 0) then ; allocate(p5(2,npp), s5(npp), o5(npp)) ; p5 = ppprs ; s5 = schwpp ; o5 = orderpp
    d%npair(5) = npp ; d%prs(5) = c_loc(p5) ; d%schw(5) = c_loc(s5) ; d%order(5) = c_loc(o5) ; d%last(5) = lastpp ; end if
  if (npd > 0) then ; allocate(p6(2,npd), s6(npd), o6(npd)) ; p6 = pdprs ; s6 = schwpd ; o6 = orderpd
    d%npair(6) = npd ; d%prs(6) = c_loc(p6) ; d%schw(6) = c_loc(s6) ; d%order(6) = c_loc(o6) ; d%last(6) = lastpd ; end if
  if (npl > 0) then ; allocate(p7(2,npl), s7(npl), o7(npl)) ; p7 = plprs ; s7 = schwpl ; o7 = orderpl
    d%npair(7) = npl ; d%prs(7) = c_loc(p7) ; d%schw(7) = c_loc(s7) ; d%order(7) = c_loc(o7) ; d%last(7) = lastpl ; end if
  if (ndd > 0) then ; allocate(p8(2,ndd), s8(ndd), o8(ndd)) ; p8 = ddprs ; s8 = schwdd ; o8 = orderdd
    d%npair(8) = ndd ; d%prs(8) = c_loc(p8) ; d%schw(8) = c_loc(s8) ; d%order(8) = c_loc(o8) ; d%last(8) = lastdd ; end if
  if (ndl > 0) then ; allocate(p9(2,ndl), s9(ndl), o9(ndl)) ; p9 = dlprs ; s9 = schwdl ; o9 = orderdl
    d%npair(9) = ndl ; d%prs(9) = c_loc(p9) ; d%schw(9) = c_loc(s9) ; d%order(9) = c_loc(o9) ; d%last(9) = lastdl ; end if
  if (nll > 0) then ; allocate(p10(2,nll), s10(nll), o10(nll)) ; p10 = llprs ; s10 = schwll ; o10 = orderll
    d%npair(10) = nll ; d%prs(10) = c_loc(p10) ; d%schw(10) = c_loc(s10) ; d%order(10) = c_loc(o10) ; d%last(10) = lastll ; end if
  rc = hfx_ctx_create(d, hfx_ctx)
  if (rc /= 0) then
    write (*,*) "hfx_ctx_create failed with code", rc
    stop
  end if
end subroutine hfx_context

subroutine hfx_build(f, p, thr, mode)
  use nrtype
  use molprops
  real(dp), intent(in) :: p(nb,nb), thr
  real(dp), intent(inout) :: f(nb,nb)
  integer(c_int), intent(in) :: mode
  integer(c_int) :: rc
  call hfx_context()
  rc = hfx_fock_nd(hfx_ctx, p, f, int(nb, c_int), real(thr, c_double), mode)
  if (rc /= 0) then
    write (*,*) "hfx_fock_nd failed with code", rc
    stop
  end if
end subroutine hfx_build

end module hfx_fortran

subroutine fock_build_non_diag(f, p, thr)
use nrtype ; use molprops ; use hfx_fortran
implicit none
real(dp), intent(in) :: p(nb,nb), thr
real(dp), intent(inout) :: f(nb,nb)
call hfx_build(f, p, thr, HFX_JK)
end subroutine fock_build_non_diag

subroutine fock_build_non_diag_j_only(f, p, thr)
use nrtype ; use molprops ; use hfx_fortran
implicit none
real(dp), intent(in) :: p(nb,nb), thr
real(dp), intent(inout) :: f(nb,nb)
call hfx_build(f, p, thr, HFX_J_ONLY)
end subroutine fock_build_non_diag_j_only
