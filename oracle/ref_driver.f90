! Driver around the reference's own BIEF (built by oracle/ref.mk into
! oracle/_ref/libbief_ref.a). It is written for this project: it fills a
! BIEF_MESH with the fields the linear-system path reads, then calls the
! reference routines SOLVE, MATVEC, DIRICH, LUMP, MT01AA, MT02AA, MT01PP,
! MT02PP, and the SUPG routines KSPG11, MT03AA, MT04AA, MT05AA unchanged, and
! runs telemac2d PROPAG's wave-equation operand sequence with the reference's
! element routines, ASSVEC, OV/OVD, MATVEC, DIRICH, SOLVE and telemac2d's
! FRICTION_CALC and FRICTI (ref_propag).
!
! TEST INFRASTRUCTURE ONLY (checker + "reference" CPU baseline). Entry points
! are bind(C) so tests/ and bench.py's cpu_baseline leg can load the shared
! library with ctypes. Connectivity arrives 0-based, structure of arrays.
module ref_driver
  use bief
  use declarations_special
  use interface_telemac2d, only: fricti, friction_calc
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

  ! One wave-equation step of telemac2d PROPAG (TREATMENT OF THE LINEAR
  ! SYSTEM = 2, propag.F:664-716, 1003-1012, 1150-1170, 1369-1465, 1634-1710)
  ! for these options: P1 depth and velocity, implicitation 1 for depth and
  ! velocity (TETAH = TETAU = TETAZCOMP = 1), no advection of velocity, no
  ! diffusion of velocity, no sources, no tidal flats (OPTBAN 0), one
  ! sub-iteration (HPROP = max(HN, 0), hpropa.f), Strickler friction (KFROT 3,
  ! HFROT 1), no vertical structures, sequential, MODASS 1, IORDRH 0 (DH
  ! starts at 0), depth prescribed (KDIR) where limdir == kdir with the depth
  ! hbor, walls elsewhere (no boundary flux term: FLBOR vanishes on wall
  ! segments, and free-exit segments join prescribed-depth points only).
  ! Every operand is computed by the reference's routine for it:
  !   VOLU2D  VC00AA + ASSVEC (VECTOR 'MASBAS')
  !   CF      OV 'X=+(Y,C)' on HN (friction_unif.f:124-127), FRICTION_CALC
  !   FU      FRICTI; SLOP10 (SLOPES) and the masked MASBAS of propag.F:1155-1161
  !   AM2%D   OV 'X=CY', 'X=X+Y' (OM 'M=M+D'), OVD 'X=1/Y' guarded at 1e-6
  !   DM1     OV 'X=CYZ' with C = -GRAV
  !   AM1     MT01AA (SL1 = 1/dt) 'M=N', MT02AA 'MATDIFUV' (-1, DM1, HPROP)
  !           'M=M+N' (assembled diagonals added, off-diagonal terms added)
  !   CV2,CV3 VC13AA + ASSVEC (VECTOR 'GRADF', -GRAV, T8 = HN + ZF),
  !           OV 'X=X+YZ' with AM3%D = VOLU2D/dt and UN, VN
  !   T2,T3   OV 'X=YZ' with the inverted AM2%D (= AM3%D)
  !   CV1     0, MATVEC 'X=X-AY' with BM1, BM2 = MT08AA (MATFGR X/Y, TETAU,
  !           HPROP) and T2, T3
  !   DIRICH  on DH with HBOR - HN; then SOLVE (CG, PRECON 2);
  !   U, V    CV2 += VC13AA(-GRAV DH) assembled (VECTOR 'GRADF' '+'), OV 'X=YZ'
  ! Outputs are node vectors (npoin) unless stated: volu2d, cf, fu, am2d (the
  ! inverted diagonal), dm1, hprop, t2, t3, da and xa(nelem, 3) of AM1 after
  ! DIRICH, cv1 after DIRICH, dh0 (DH after DIRICH), dh (after SOLVE), u, v.
  subroutine ref_propag(nelem, npoin, ikle0, xel, yel, surfac, hn, un, vn, zf, chestr, hbor, &
       limdir, kdir, dt, grav, nitmax, eps, volu2d, cf, fu, am2d, dm1, hprop, t2, t3, da, &
       xa, cv1, dh0, dh, u, v, niter) bind(c, name='ref_propag')
    integer(c_int), value :: nelem, npoin, kdir, nitmax
    integer(c_int), intent(in) :: ikle0(nelem, 3), limdir(*)
    real(c_double), intent(in) :: xel(nelem, 3), yel(nelem, 3), surfac(nelem)
    real(c_double), intent(in) :: hn(npoin), un(npoin), vn(npoin), zf(npoin), chestr(npoin)
    real(c_double), intent(in) :: hbor(npoin)
    real(c_double), value :: dt, grav, eps
    real(c_double), intent(out) :: volu2d(npoin), cf(npoin), fu(npoin), am2d(npoin)
    real(c_double), intent(out) :: dm1(npoin), hprop(npoin), t2(npoin), t3(npoin)
    real(c_double), intent(out) :: da(npoin), xa(nelem, 3), cv1(npoin), dh0(npoin)
    real(c_double), intent(out) :: dh(npoin), u(npoin), v(npoin)
    integer(c_int), intent(out) :: niter
    integer, allocatable :: ik(:,:)
    real(c_double), allocatable :: w(:,:), te3(:), tm(:), am3d(:), d1(:), d2(:)
    real(c_double), allocatable :: tmass(:,:), xmass(:,:), tdif(:,:), xdif(:,:)
    real(c_double), allocatable :: a(:,:), t8(:), cv2(:), cv3(:), hc(:)
    real(c_double), allocatable :: xbm(:,:), dbm(:)
    type(bief_obj) :: sf, sdm1, shp, maskpt
    type(bief_obj), save :: oun, ovn, ohn, ocf, ofu, ofv, ofud, ofvd, ot1, ot2, ounsv, omsk, &
         ochestr, ohc, ovres
    type(slvcfg) :: cfg
    real(c_double) :: dummy(1)
    integer :: icoord
    character(len=16) :: formul
    allocate(ik(nelem, 3), w(nelem, 3), te3(nelem), tm(npoin), am3d(npoin), d1(npoin), d2(npoin))
    allocate(tmass(nelem, 3), xmass(nelem, 3), tdif(nelem, 3), xdif(nelem, 3), a(nelem, 9))
    allocate(t8(npoin), cv2(npoin), cv3(npoin), hc(npoin), xbm(nelem, 6), dbm(npoin))
    ik = ikle0 + 1
    sf%elm = 11
    sf%type = 2
    sf%dim2 = 1
    ! VOLU2D: VECTOR 'MASBAS' (vc00aa.f + assvec.f)
    call vc00aa(1.d0, surfac, nelem, nelem, w(1,1), w(1,2), w(1,3))
    call assvec(volu2d, ik, npoin, nelem, nelem, 11, w, .true., 1, .false., dummy, 3)
    ! friction coefficient (friction_unif.f:120-146; friction_calc.f:130-134)
    call bief_allvec(1, ochestr, 'CHESTR', 11, 1, 2, mesh)
    call bief_allvec(1, ohc, 'HC    ', 11, 1, 2, mesh)
    call bief_allvec(1, ovres, 'VRES  ', 11, 1, 2, mesh)
    call bief_allvec(1, ocf, 'CF    ', 11, 1, 2, mesh)
    ochestr%r(1:npoin) = chestr
    call ov('X=Y     ', hc, hn, hn, 0.d0, npoin)
    call ov('X=+(Y,C)', hc, hc, hc, 1.d-4, npoin)
    ohc%r(1:npoin) = hc
    ovres%r = 0.d0
    call friction_calc(1, npoin, 3, 0.d0, 0.d0, grav, 0.4d0, ochestr, ohc, ohc, ovres, ocf)
    cf = ocf%r(1:npoin)
    ! implicit friction FU_IMP (fricti.f, HFROT 1)
    call bief_allvec(1, oun, 'UN    ', 11, 1, 2, mesh)
    call bief_allvec(1, ovn, 'VN    ', 11, 1, 2, mesh)
    call bief_allvec(1, ohn, 'HN    ', 11, 1, 2, mesh)
    call bief_allvec(1, ofu, 'FU    ', 11, 1, 2, mesh)
    call bief_allvec(1, ofv, 'FV    ', 11, 1, 2, mesh)
    call bief_allvec(1, ofud, 'FUD   ', 11, 1, 2, mesh)
    call bief_allvec(1, ofvd, 'FVD   ', 11, 1, 2, mesh)
    call bief_allvec(1, ot1, 'T1    ', 11, 1, 2, mesh)
    call bief_allvec(1, ot2, 'T2    ', 11, 1, 2, mesh)
    call bief_allvec(1, ounsv, 'UNSV2D', 11, 1, 2, mesh)
    call bief_allvec(1, omsk, 'MASKEL', 10, 1, 2, mesh)
    oun%r(1:npoin) = un
    ovn%r(1:npoin) = vn
    ohn%r(1:npoin) = hn
    call fricti(ofu, ofv, ofud, ofvd, oun, ovn, ohn, ocf, mesh, ot1, ot2, .false., ounsv, &
         .false., omsk, 1)
    fu = ofu%r(1:npoin)
    ! T2 = -MASBAS with the slope mask TE3 (SLOPES), T3 = FU T2 (propag.F:1155-1161)
    call slop10(te3, xel, yel, zf, ik, nelem, nelem)
    call vc00aa(-1.d0, surfac, nelem, nelem, w(1,1), w(1,2), w(1,3))
    call assvec(tm, ik, npoin, nelem, nelem, 11, w, .true., 1, .true., te3, 3)
    call ov('X=XY    ', tm, fu, fu, 0.d0, npoin)
    ! AM2%D = VOLU2D / DT (propag.F:707), AM3%D the same, + friction (1169)
    call ov('X=CY    ', am2d, volu2d, volu2d, 1.d0 / dt, npoin)
    am3d = am2d
    call ov('X=X+Y   ', am2d, tm, tm, 0.d0, npoin)
    ! CV2, CV3 (propag.F:581, 602-606, 1010-1011), before the friction is added to AM3
    call ov('X=Y+Z   ', t8, hn, zf, 0.d0, npoin)
    do icoord = 1, 2
      call vc13aa(-grav, sf, t8, xel, yel, ik(1,1), ik(1,2), ik(1,3), ik(1,1), nelem, nelem, &
           w(1,1), w(1,2), w(1,3), icoord)
      if (icoord == 1) then
        call assvec(cv2, ik, npoin, nelem, nelem, 11, w, .true., 1, .false., dummy, 3)
        call ov('X=X+YZ  ', cv2, am3d, un, 0.d0, npoin)
      else
        call assvec(cv3, ik, npoin, nelem, nelem, 11, w, .true., 1, .false., dummy, 3)
        call ov('X=X+YZ  ', cv3, am3d, vn, 0.d0, npoin)
      end if
    end do
    ! inversion (propag.F:1384) and DM1 (1397)
    call ovd('X=1/Y   ', am2d, am2d, am2d, 0.d0, npoin, 2, 0.d0, 1.d-6)
    call ov('X=CYZ   ', dm1, volu2d, am2d, -grav, npoin)
    ! HPROP = max(HN, 0) (hpropa.f)
    call ov('X=Y     ', hprop, hn, hn, 0.d0, npoin)
    call ov('X=+(Y,C)', hprop, hprop, hprop, 0.d0, npoin)
    ! AM1 = MATMAS / DT (propag.F:679) + MATDIFUV (-1, DM1, HPROP) (1407)
    call mt01aa(tmass(1,1), xmass(1,1), xmass(1,2), tmass(1,2), xmass(1,3), tmass(1,3), &
         1.d0 / dt, surfac, nelem, nelem)
    sdm1%elm = 11
    sdm1%type = 2
    sdm1%dim2 = 1
    sdm1%maxdim1 = npoin
    shp = sdm1
    formul = 'MATDIFUV        '
    call mt02aa(tdif(1,1), xdif(1,1), xdif(1,2), tdif(1,2), xdif(1,3), tdif(1,3), -1.d0, &
         sdm1, dm1, shp, hprop, xel, yel, surfac, ik(1,1), ik(1,2), ik(1,3), nelem, nelem, &
         formul)
    call assvec(d1, ik, npoin, nelem, nelem, 11, tmass, .true., 1, .false., dummy, 3)
    call assvec(d2, ik, npoin, nelem, nelem, 11, tdif, .true., 1, .false., dummy, 3)
    call ov('X=Y+Z   ', da, d1, d2, 0.d0, npoin)
    xa = xmass + xdif
    ! T2, T3 (propag.F:1413-1414)
    call ov('X=YZ    ', t2, cv2, am2d, 0.d0, npoin)
    call ov('X=YZ    ', t3, cv3, am2d, 0.d0, npoin)
    ! CV1 = - BM1 T2 - BM2 T3 (propag.F:745-754, 466-499, 1464-1465)
    cv1 = 0.d0
    do icoord = 1, 2
      call mt08aa(a(1,1), a(1,2), a(1,3), a(1,4), a(1,5), a(1,6), a(1,7), a(1,8), a(1,9), &
           1.d0, sf, hprop, xel, yel, ik(1,1), ik(1,2), ik(1,3), nelem, nelem, icoord)
      ! the diagonal assembled (A11, A22, A33), the off-diagonal terms in the
      ! 'Q' order 12, 13, 23, 21, 31, 32
      w(:,1) = a(:,1)
      w(:,2) = a(:,5)
      w(:,3) = a(:,9)
      call assvec(dbm, ik, npoin, nelem, nelem, 11, w, .true., 1, .false., dummy, 3)
      xbm(:,1) = a(:,2)
      xbm(:,2) = a(:,3)
      xbm(:,3) = a(:,6)
      xbm(:,4) = a(:,4)
      xbm(:,5) = a(:,7)
      xbm(:,6) = a(:,8)
      call load_matrix('Q', 'Q', dbm, xbm)
      vx%r(1:npoin) = cv1
      if (icoord == 1) then
        vy%r(1:npoin) = t2
      else
        vy%r(1:npoin) = t3
      end if
      call matvec('X=X-AY  ', vx, ma, vy, 0.d0, mesh)
      cv1 = vx%r(1:npoin)
    end do
    ! DIRICH on DH with HBOR - HN (propag.F:462, 1647), DH = 0 first (IORDRH 0)
    call load_matrix('Q', 'S', da, xa)
    vb%r(1:npoin) = cv1
    vf%r(1:npoin) = 0.d0
    ! HBOR - HN on the prescribed points (OSBD, propag.F:462): a domain-sized
    ! FBOR must be 0 elsewhere (diri01.f:85-87; a boundary-sized HBOR gets
    ! the same through OSDBIF)
    vfbor%r(1:npoin) = hbor
    call ov('X=X-Y   ', vfbor%r, hn, hn, 0.d0, npoin)
    where (limdir(1:npoin) /= kdir) vfbor%r(1:npoin) = 0.d0
    call dirich(vf, ma, vb, vfbor, limdir, vwork, mesh, kdir, .false., maskpt)
    cv1 = vb%r(1:npoin)
    dh0 = vf%r(1:npoin)
    da = ma%d%r(1:npoin)
    xa = reshape(ma%x%r(1:nelem * 3), (/nelem, 3/))
    ! SOLVE (CG, diagonal preconditioning)
    cfg%slv = 1
    cfg%precon = 2
    cfg%krylov = 5
    cfg%nitmax = nitmax
    cfg%eps = eps
    cfg%zero = 1.d-12
    cfg%ok = .true.
    cfg%nit = 0
    vx%r(1:npoin) = dh0
    call solve(vx, ma, vb, tb, cfg, .false., mesh, ma)
    dh = vx%r(1:npoin)
    niter = cfg%nit
    ! momentum (propag.F:1685-1707): CV2 += GRADF(-GRAV DH), U = CV2 AM2%D
    do icoord = 1, 2
      call vc13aa(-grav, sf, dh, xel, yel, ik(1,1), ik(1,2), ik(1,3), ik(1,1), nelem, nelem, &
           w(1,1), w(1,2), w(1,3), icoord)
      call assvec(tm, ik, npoin, nelem, nelem, 11, w, .true., 1, .false., dummy, 3)
      if (icoord == 1) then
        call ov('X=X+Y   ', cv2, tm, tm, 0.d0, npoin)
        call ov('X=YZ    ', u, cv2, am2d, 0.d0, npoin)
      else
        call ov('X=X+Y   ', cv3, tm, tm, 0.d0, npoin)
        call ov('X=YZ    ', v, cv3, am2d, 0.d0, npoin)
      end if
    end do
    deallocate(ik, w, te3, tm, am3d, d1, d2, tmass, xmass, tdif, xdif, a, t8, cv2, cv3, hc)
    deallocate(xbm, dbm)
  end subroutine

end module ref_driver
