! Driver around the reference's own BIEF (built by oracle/ref.mk into
! oracle/_ref/libbief_ref.a). It is written for this project: it fills a
! BIEF_MESH with the fields the linear-system path reads, then calls the
! reference routines SOLVE, MATVEC, DIRICH, LUMP, MT01AA, MT02AA, MT01PP,
! MT02PP, and the SUPG routines KSPG11, MT03AA, MT04AA, MT05AA unchanged.
!
! TEST INFRASTRUCTURE ONLY (checker + "reference" CPU baseline). Entry points
! are bind(C) so tests/ and bench.py's cpu_baseline leg can load the shared
! library with ctypes. Connectivity arrives 0-based, structure of arrays.
module ref_driver
  use bief
  use declarations_special
  use iso_c_binding
  implicit none
  type(bief_mesh), target, save :: mesh
  type(bief_obj), target, save :: ma, vx, vy, vb, tb, vf, vfbor, vwork
  integer, save :: npoin_s = 0, nelem_s = 0, ndp_s = 3, ielm_s = 11
  logical, save :: ready = .false.
  character(len=8), parameter :: opnames(0:10) = (/ &
       'X=AY    ', 'X=CAY   ', 'X=-AY   ', 'X=X+AY  ', 'X=X-AY  ', &
       'X=X+CAY ', 'X=TAY   ', 'X=-TAY  ', 'X=X+TAY ', 'X=X-TAY ', &
       'X=X+CTAY' /)
contains

  subroutine set_int_obj(obj, name, n1, n2)
    type(bief_obj), intent(inout) :: obj
    character(len=6), intent(in) :: name
    integer, intent(in) :: n1, n2
    obj%key = 123456
    obj%type = 2
    obj%nat = 2
    obj%name = name
    obj%father = 'XXXXXX'
    obj%elm = 10
    obj%dim1 = n1
    obj%maxdim1 = n1
    obj%dim2 = n2
    obj%maxdim2 = n2
    obj%dimdisc = 0
    obj%status = 1
    obj%typr = '?'
    obj%typi = '?'
    allocate(obj%i(max(n1*n2, 1)))
    obj%i = 0
    nullify(obj%r)
    nullify(obj%e)
  end subroutine

  subroutine ref_mesh_init(ndp, npoin2, nelem2, nplan, ikle0) bind(c, name='ref_mesh_init')
    integer(c_int), value :: ndp, npoin2, nelem2, nplan
    integer(c_int), intent(in) :: ikle0(*)
    integer :: npoin, nelem, k, e, cfgm(2)
    if (ready) call ref_mesh_free_all()
    nullify(mesh%w, mesh%t)
    lng = 2
    lu = 6
    ncsize = 0
    if (ndp == 3) then
      npoin = npoin2
      nelem = nelem2
      ielm_s = 11
    else
      npoin = npoin2 * nplan
      nelem = nelem2 * (nplan - 1)
      ielm_s = 41
    end if
    npoin_s = npoin
    nelem_s = nelem
    ndp_s = ndp
    mesh%name = 'MESH  '
    allocate(mesh%nelem, mesh%nelmax, mesh%nptfr, mesh%nptfrx, mesh%neleb, &
         mesh%nelebx, mesh%nseg, mesh%nsegbor, mesh%dim1, mesh%typelm, &
         mesh%typelmbnd, mesh%npoin, mesh%npmax, mesh%mxptvs, mesh%mxelvs, &
         mesh%lv)
    mesh%nelem = nelem
    mesh%nelmax = nelem
    mesh%nptfr = 1
    mesh%nptfrx = 1
    mesh%neleb = 0
    mesh%nelebx = 0
    mesh%nseg = 0
    mesh%nsegbor = 0
    mesh%npoin = npoin
    mesh%npmax = npoin
    mesh%mxptvs = 1
    mesh%mxelvs = 1
    mesh%lv = 1
    if (ndp == 3) then
      mesh%dim1 = 2
      mesh%typelm = 10
    else
      mesh%dim1 = 3
      mesh%typelm = 40
    end if
    mesh%typelmbnd = 1
    allocate(mesh%nds(0:81, 7))
    mesh%nds = 0
    call bief_ininds(npoin2, 1, nelem2, npoin2, 1, nelem2, nplan, 0, mesh%nds, 0)
    allocate(mesh%ikle, mesh%nbor, mesh%iklem1, mesh%limvoi, mesh%gloseg, mesh%ifac)
    call set_int_obj(mesh%ikle, 'IKLE  ', nelem, ndp)
    do k = 1, ndp
      do e = 1, nelem
        mesh%ikle%i((k-1)*nelem + e) = ikle0((k-1)*nelem + e) + 1
      end do
    end do
    call set_int_obj(mesh%nbor, 'NBOR  ', 1, 1)
    call set_int_obj(mesh%iklem1, 'IKLEM1', 1, 1)
    call set_int_obj(mesh%limvoi, 'LIMVOI', 1, 1)
    call set_int_obj(mesh%gloseg, 'GLOSEG', 1, 2)
    call set_int_obj(mesh%ifac, 'IFAC  ', npoin, 1)
    mesh%ifac%i = 1
    allocate(mesh%w, mesh%t)
    ! status 0: the "element" argument is the first dimension itself
    call bief_allvec(1, mesh%w, 'W     ', nelem, ndp, 0, mesh)
    call bief_allvec(1, mesh%t, 'T     ', ielm_s, 1, 2, mesh)
    call bief_allvec(1, vx, 'X     ', ielm_s, 1, 2, mesh)
    call bief_allvec(1, vy, 'Y     ', ielm_s, 1, 2, mesh)
    call bief_allvec(1, vb, 'B     ', ielm_s, 1, 2, mesh)
    call bief_allvec(1, vf, 'F     ', ielm_s, 1, 2, mesh)
    call bief_allvec(1, vfbor, 'FBOR  ', ielm_s, 1, 2, mesh)
    call allblo(vwork, 'WORK  ')
    call bief_allvec_in_block(vwork, 2, 1, 'WRK   ', ielm_s, 1, 2, mesh)
    call allblo(tb, 'TB    ')
    call bief_allvec_in_block(tb, 50, 1, 'TBV   ', ielm_s, 1, 2, mesh)
    cfgm(1) = 1
    cfgm(2) = 1
    call bief_allmat(ma, 'A     ', ielm_s, ielm_s, cfgm, 'Q', 'Q', mesh)
    ! XA(NELMAX,*) layout of the off-diagonal terms (matvct.f:472)
    ma%stox = 1
    ready = .true.
  end subroutine

  ! Segment storage (MATRIX STORAGE = 3) on the current triangle mesh:
  ! GLOSEG (nseg, 2) arrives 0-based, structure of arrays; the driver's
  ! matrix is re-allocated with storage 3 and the load/store helpers then
  ! treat the segments as two-node elements.
  subroutine ref_mesh_segments(nseg, gloseg0) bind(c, name='ref_mesh_segments')
    integer(c_int), value :: nseg
    integer(c_int), intent(in) :: gloseg0(*)
    integer :: cfgm(2)
    mesh%nseg = nseg
    mesh%nds(11, 2) = nseg
    deallocate(mesh%gloseg%i)
    call set_int_obj_alloc(mesh%gloseg, nseg, 2)
    mesh%gloseg%i(1:2 * nseg) = gloseg0(1:2 * nseg) + 1
    nullify(ma%d, ma%x)
    cfgm(1) = 3
    cfgm(2) = 1
    call bief_allmat(ma, 'A     ', ielm_s, ielm_s, cfgm, 'Q', 'Q', mesh)
    ndp_s = 2
    nelem_s = nseg
  end subroutine

  subroutine set_int_obj_alloc(obj, n1, n2)
    type(bief_obj), intent(inout) :: obj
    integer, intent(in) :: n1, n2
    obj%dim1 = n1
    obj%maxdim1 = n1
    obj%dim2 = n2
    obj%maxdim2 = n2
    allocate(obj%i(max(n1 * n2, 1)))
    obj%i = 0
  end subroutine

  subroutine ref_mesh_free_all()
    ! objects of the previous mesh are dropped (not freed: this is a test
    ! driver and meshes are few) so BIEF allocates them afresh
    nullify(vx%r, vy%r, vb%r, vf%r, vfbor%r)
    nullify(vwork%adr, tb%adr, ma%d, ma%x)
    vwork%n = 0
    tb%n = 0
    ready = .false.
  end subroutine

  subroutine load_matrix(typdia, typext, da, xa)
    character(kind=c_char), value :: typdia, typext
    real(c_double), intent(in) :: da(*), xa(*)
    integer :: ncoef, npair
    npair = ndp_s * (ndp_s - 1) / 2
    ncoef = npair
    if (typext == 'Q') ncoef = 2 * npair
    ma%typdia = typdia
    ma%typext = typext
    ma%d%r(1:npoin_s) = da(1:npoin_s)
    if (typext /= '0') ma%x%r(1:nelem_s * ncoef) = xa(1:nelem_s * ncoef)
  end subroutine

  subroutine store_matrix(typdia, da, xa)
    character(kind=c_char), intent(out) :: typdia
    real(c_double), intent(inout) :: da(*), xa(*)
    integer :: ncoef, npair
    npair = ndp_s * (ndp_s - 1) / 2
    ncoef = npair
    if (ma%typext == 'Q') ncoef = 2 * npair
    typdia = ma%typdia
    da(1:npoin_s) = ma%d%r(1:npoin_s)
    if (ma%typext /= '0') xa(1:nelem_s * ncoef) = ma%x%r(1:nelem_s * ncoef)
  end subroutine

  subroutine ref_matvec(op, typdia, typext, da, xa, y, c, x) bind(c, name='ref_matvec')
    integer(c_int), value :: op
    character(kind=c_char), value :: typdia, typext
    real(c_double), intent(in) :: da(*), xa(*), y(*)
    real(c_double), value :: c
    real(c_double), intent(inout) :: x(*)
    call load_matrix(typdia, typext, da, xa)
    vy%r(1:npoin_s) = y(1:npoin_s)
    vx%r(1:npoin_s) = x(1:npoin_s)
    call matvec(opnames(op), vx, ma, vy, c, mesh)
    x(1:npoin_s) = vx%r(1:npoin_s)
  end subroutine

  subroutine ref_solve(slv, precon, krylov, nitmax, eps, typdia, typext, da, xa, &
       b, x, logpath, loglen) bind(c, name='ref_solve')
    integer(c_int), value :: slv, precon, krylov, nitmax, loglen
    real(c_double), value :: eps
    character(kind=c_char), intent(inout) :: typdia
    character(kind=c_char), value :: typext
    real(c_double), intent(inout) :: da(*), xa(*), b(*), x(*)
    character(kind=c_char), intent(in) :: logpath(*)
    type(slvcfg) :: cfg
    type(bief_obj) :: aux
    character(len=512) :: path
    integer :: i, oldlu
    path = ' '
    do i = 1, min(loglen, 512)
      path(i:i) = logpath(i)
    end do
    call load_matrix(typdia, typext, da, xa)
    vb%r(1:npoin_s) = b(1:npoin_s)
    vx%r(1:npoin_s) = x(1:npoin_s)
    cfg%slv = slv
    cfg%precon = precon
    cfg%krylov = krylov
    cfg%nitmax = nitmax
    cfg%eps = eps
    cfg%zero = 1.d-12
    cfg%ok = .true.
    cfg%nit = 0
    oldlu = lu
    if (loglen > 0) then
      open(unit=91, file=trim(path), status='replace', action='write')
      lu = 91
    end if
    call solve(vx, ma, vb, tb, cfg, .true., mesh, ma)
    if (loglen > 0) then
      close(91)
      lu = oldlu
    end if
    x(1:npoin_s) = vx%r(1:npoin_s)
    b(1:npoin_s) = vb%r(1:npoin_s)
    call store_matrix(typdia, da, xa)
  end subroutine

  subroutine ref_dirich(typdia, typext, da, xa, sm, fbor, limdir, kdir, f) bind(c, name='ref_dirich')
    character(kind=c_char), value :: typdia, typext
    real(c_double), intent(inout) :: da(*), xa(*), sm(*), f(*)
    real(c_double), intent(in) :: fbor(*)
    integer(c_int), intent(in) :: limdir(*)
    integer(c_int), value :: kdir
    character(kind=c_char) :: td
    type(bief_obj) :: maskpt
    call load_matrix(typdia, typext, da, xa)
    vb%r(1:npoin_s) = sm(1:npoin_s)
    vf%r(1:npoin_s) = f(1:npoin_s)
    vfbor%r(1:npoin_s) = fbor(1:npoin_s)
    call dirich(vf, ma, vb, vfbor, limdir, vwork, mesh, kdir, .false., maskpt)
    sm(1:npoin_s) = vb%r(1:npoin_s)
    f(1:npoin_s) = vf%r(1:npoin_s)
    call store_matrix(td, da, xa)
  end subroutine

  subroutine ref_lump(typdia, typext, da, xa, xmul, diag) bind(c, name='ref_lump')
    character(kind=c_char), value :: typdia, typext
    real(c_double), intent(in) :: da(*), xa(*)
    real(c_double), value :: xmul
    real(c_double), intent(out) :: diag(*)
    call load_matrix(typdia, typext, da, xa)
    call lump(vx, ma, mesh, xmul)
    diag(1:npoin_s) = vx%r(1:npoin_s)
  end subroutine

  ! element matrices, called directly (arrays only)
  subroutine ref_mt01aa(nelem, xmul, surfac, t, xm) bind(c, name='ref_mt01aa')
    integer(c_int), value :: nelem
    real(c_double), value :: xmul
    real(c_double), intent(in) :: surfac(*)
    real(c_double), intent(out) :: t(nelem, 3), xm(nelem, 3)
    call mt01aa(t(1,1), xm(1,1), xm(1,2), t(1,2), xm(1,3), t(1,3), xmul, surfac, nelem, nelem)
  end subroutine

  subroutine ref_mt02aa(kind, nelem, npoin, xmul, u, v, xel, yel, surfac, ikle0, mon, t, xm) &
       bind(c, name='ref_mt02aa')
    integer(c_int), value :: kind, nelem, npoin, mon
    real(c_double), value :: xmul
    real(c_double), intent(in) :: u(*), v(*), xel(*), yel(*), surfac(*)
    integer(c_int), intent(in) :: ikle0(nelem, 3)
    real(c_double), intent(out) :: t(nelem, 3), xm(nelem, 3)
    type(bief_obj) :: su, sv
    character(len=16) :: formul
    integer, allocatable :: ik(:,:)
    allocate(ik(nelem, 3))
    ik = ikle0 + 1
    formul = 'MATDIF          '
    if (kind == 2) formul = 'MATDIFUV        '
    if (mon /= 0) formul(14:16) = 'MON'
    su%type = 2
    sv%type = 2
    su%dim2 = 1
    sv%dim2 = 1
    su%elm = 11
    sv%elm = 11
    su%maxdim1 = npoin
    if (kind == 0) su%elm = 10
    if (kind == 3) then
      su%elm = 10
      su%dim2 = 3
    end if
    if (kind == 4) su%dim2 = 3
    call mt02aa(t(1,1), xm(1,1), xm(1,2), t(1,2), xm(1,3), t(1,3), xmul, su, u, sv, v, &
         xel, yel, surfac, ik(1,1), ik(1,2), ik(1,3), nelem, nelem, formul)
    deallocate(ik)
  end subroutine

  subroutine ref_mt01pp(nelem, xmul, z, surfac, ikle0, t, xm) bind(c, name='ref_mt01pp')
    integer(c_int), value :: nelem
    real(c_double), value :: xmul
    real(c_double), intent(in) :: z(*), surfac(*)
    integer(c_int), intent(in) :: ikle0(nelem, 6)
    real(c_double), intent(out) :: t(nelem, 6), xm(nelem, 30)
    integer, allocatable :: ik(:,:)
    allocate(ik(nelem, 6))
    ik = ikle0 + 1
    call mt01pp(t, xm, xmul, z, surfac, ik, nelem, nelem)
    deallocate(ik)
  end subroutine

  ! MT02PP with P1 coefficients f, g, h on the prism nodes (npoin of them);
  ! h_zero: H of type '0'; p0_vertical: H%DIMDISC = 4111
  subroutine ref_mt02pp(nelem, npoin, nplan, xmul, f, g, h, h_zero, p0_vertical, xel, yel, z, &
       surfac, ikle0, inchyd, mon, t, xm) bind(c, name='ref_mt02pp')
    integer(c_int), value :: nelem, npoin, nplan, h_zero, p0_vertical, inchyd, mon
    real(c_double), value :: xmul
    real(c_double), intent(in) :: f(*), g(*), h(*), z(*), surfac(*)
    real(c_double), intent(in) :: xel(nelem, 3), yel(nelem, 3)
    integer(c_int), intent(in) :: ikle0(nelem, 6)
    real(c_double), intent(out) :: t(nelem, 6), xm(nelem, 15)
    type(bief_obj) :: sf, sg, sh
    integer, allocatable :: ik(:,:)
    double precision, allocatable :: x6(:,:), y6(:,:)
    character(len=16) :: formul
    allocate(ik(nelem, 6), x6(nelem, 6), y6(nelem, 6))
    ik = ikle0 + 1
    x6 = 0.d0
    y6 = 0.d0
    x6(:, 1:3) = xel
    y6(:, 1:3) = yel
    x6(:, 4:6) = xel
    y6(:, 4:6) = yel
    sf%elm = 41
    sf%typr = 'Q'
    sf%dimdisc = 0
    sf%dim1 = npoin
    sg = sf
    sh = sf
    if (h_zero /= 0) sh%typr = '0'
    if (p0_vertical /= 0) sh%dimdisc = 4111
    formul = 'MATDIF          '
    if (mon /= 0) formul(14:16) = 'MON'
    call mt02pp(t, xm, xmul, sf, sg, sh, f, g, h, x6, y6, z, surfac, ik, nelem, nelem, &
         inchyd /= 0, formul, nplan)
    deallocate(ik, x6, y6)
  end subroutine

  ! KSPG11: SUPG element vectors (P0) from P1 velocities
  subroutine ref_kspg11(nelem, xmul, u, v, xel, yel, ikle0, kx, ky) bind(c, name='ref_kspg11')
    integer(c_int), value :: nelem
    real(c_double), value :: xmul
    real(c_double), intent(in) :: u(*), v(*), xel(nelem, 3), yel(nelem, 3)
    integer(c_int), intent(in) :: ikle0(nelem, 3)
    real(c_double), intent(out) :: kx(nelem), ky(nelem)
    integer, allocatable :: ik(:,:)
    allocate(ik(nelem, 3))
    ik = ikle0 + 1
    call kspg11(kx, ky, xel, yel, u, v, ik, nelem, nelem, xmul)
    deallocate(ik)
  end subroutine

  ! MT03AA (MASUPG) with P0 F, G and P1 U, V; xm in the 'Q' order 12,13,23,21,31,32
  subroutine ref_mt03aa(nelem, xmul, f, g, u, v, xel, yel, surfac, ikle0, t, xm) &
       bind(c, name='ref_mt03aa')
    integer(c_int), value :: nelem
    real(c_double), value :: xmul
    real(c_double), intent(in) :: f(*), g(*), u(*), v(*), xel(nelem, 3), yel(nelem, 3)
    real(c_double), intent(in) :: surfac(*)
    integer(c_int), intent(in) :: ikle0(nelem, 3)
    real(c_double), intent(out) :: t(nelem, 3), xm(nelem, 6)
    type(bief_obj) :: sf, sg, su, sv
    integer, allocatable :: ik(:,:)
    allocate(ik(nelem, 4))
    ik(:, 1:3) = ikle0 + 1
    ik(:, 4) = ik(:, 1)
    sf%elm = 10
    sg%elm = 10
    su%elm = 11
    sv%elm = 11
    call mt03aa(t(1,1), xm(1,1), xm(1,2), xm(1,4), t(1,2), xm(1,3), xm(1,5), xm(1,6), &
         t(1,3), xmul, sf, sg, su, sv, f, g, u, v, xel, yel, surfac, &
         ik(1,1), ik(1,2), ik(1,3), ik(1,4), nelem, nelem)
    deallocate(ik)
  end subroutine

  ! MT04AA (MAUGUG) with P1 U, V; symmetric xm 12,13,23
  subroutine ref_mt04aa(nelem, xmul, u, v, xel, yel, surfac, ikle0, t, xm) &
       bind(c, name='ref_mt04aa')
    integer(c_int), value :: nelem
    real(c_double), value :: xmul
    real(c_double), intent(in) :: u(*), v(*), xel(nelem, 3), yel(nelem, 3), surfac(*)
    integer(c_int), intent(in) :: ikle0(nelem, 3)
    real(c_double), intent(out) :: t(nelem, 3), xm(nelem, 3)
    type(bief_obj) :: su, sv
    integer, allocatable :: ik(:,:)
    allocate(ik(nelem, 3))
    ik = ikle0 + 1
    su%elm = 11
    sv%elm = 11
    call mt04aa(t(1,1), xm(1,1), xm(1,2), t(1,2), xm(1,3), t(1,3), xmul, su, sv, u, v, &
         xel, yel, surfac, ik, nelem, nelem)
    deallocate(ik)
  end subroutine

  ! MT05AA (MATVGR): u_elm 11 (P1 U, V) or 10 (P0 U, V); upwind /= 0 selects
  ! the FORMUL(16:16) = 'N' variant; xm in the 'Q' order
  subroutine ref_mt05aa(nelem, xmul, u_elm, upwind, u, v, xel, yel, ikle0, t, xm) &
       bind(c, name='ref_mt05aa')
    integer(c_int), value :: nelem, u_elm, upwind
    real(c_double), value :: xmul
    real(c_double), intent(in) :: u(*), v(*), xel(nelem, 3), yel(nelem, 3)
    integer(c_int), intent(in) :: ikle0(nelem, 3)
    real(c_double), intent(out) :: t(nelem, 3), xm(nelem, 6)
    type(bief_obj) :: su, sv
    integer, allocatable :: ik(:,:)
    character(len=16) :: formul
    allocate(ik(nelem, 3))
    ik = ikle0 + 1
    su%elm = u_elm
    sv%elm = u_elm
    formul = 'MATVGR          '
    if (upwind /= 0) formul(16:16) = 'N'
    call mt05aa(t(1,1), xm(1,1), xm(1,2), xm(1,4), t(1,2), xm(1,3), xm(1,5), xm(1,6), &
         t(1,3), xmul, su, sv, u, v, xel, yel, ik, nelem, nelem, formul)
    deallocate(ik)
  end subroutine

end module ref_driver
