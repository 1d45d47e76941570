! instantonmod_rp.f90 -- `module instantonmod` with the ring-polymer instanton routines
! of the reference (instantonmod.f90) evaluated on the GPU, and its host-side helpers.
!
! Same module variables (pi, beta, betan, UMtilde, xtilde, well1, well2, mass, fixedends)
! and the routines the reference's drivers call (rpi_ser.f90, rpi_par.f90, pimd_par.f90):
!   on the device:  UM, UMprime, UMforceenergy, UMhessian, instanton, detJ
!   on the host:    get_align, align_atoms, rotate_vec, rotate_atoms (:222-376),
!                   spline, splint, splin_grad, locate (:397-603), findmiddle, eucliddist,
!                   read_path (:832-1035); read_hess (:1039-1074) only stops, as readhess
!                   is not supported
! The potential is reached through `module rp_bridge` (a device plugin, or any host plugin
! of the reference's API).  Chains use the reference layout x(n, ndim, natom); n, ndim,
! natom, ndof, totdof come from mcmod_mass.
!
! detJ's eigenvalues: the reference diagonalises the band matrix (DSBEVD) and its drivers
! use only sums of ln(eta) over eta > 0 and the lowest eigenvalue (rpi_ser.f90:221-237,
! 354-365).  Here no eigen-decomposition is formed: the k lowest eigenvalues (every
! negative one and the lowest positive one) are located exactly by bisection on the
! inertia, and the remaining totdof - k entries are set to the geometric mean of the
! remaining positive eigenvalues, from the block-LU log-determinant.  The array is
! ascending and every sum of ln(eta) the drivers form comes out as from the full spectrum.
module instantonmod
  use iso_c_binding
  use rp_hip
  use mcmod_mass
  use rp_bridge
  implicit none
  double precision, parameter::    pi=3.14159265358979d0
  double precision::               beta, betan, UMtilde
  double precision, allocatable::  xtilde(:,:,:), well1(:,:), well2(:,:), mass(:)
  logical::                        fixedends
  private :: ring_open, ring_close, rot_pair

contains

  ! one-chain ring handle for the current module state; a, b default to zero ends
  function ring_open(a, b) result(r)
    double precision, intent(in), optional:: a(:,:), b(:,:)
    type(c_ptr):: r
    double precision:: aa(ndim, natom), bb(ndim, natom), bn(1)
    integer(c_int):: nn(1)
    aa = 0.0d0
    bb = 0.0d0
    if (present(a)) aa = a
    if (present(b)) bb = b
    nn(1) = int(n, c_int)
    bn(1) = betan
    call rp_bridge_sync_v0()
    call rp_check(rp_ring_create(rp_bridge_pes(), mass, 1_c_int, nn, bn, aa, bb, &
         merge(1_c_int, 0_c_int, fixedends), r), 'instantonmod')
  end function ring_open

  subroutine ring_close(r)
    type(c_ptr), intent(inout):: r
    call rp_check(rp_ring_destroy(r), 'instantonmod')
    r = c_null_ptr
    call rp_bridge_restore_v0()
  end subroutine ring_close

  ! U_M of instantonmod.f90:17-46
  function UM(x, a, b)
    double precision, intent(in):: x(:,:,:)
    double precision, intent(in), optional:: a(:,:), b(:,:)
    double precision:: UM
    double precision, target:: e(1)
    type(c_ptr):: r
    r = ring_open(a, b)
    call rp_check(rp_ring_set_x(r, x), 'UM')
    call rp_check(rp_ring_um(r, c_loc(e), c_null_ptr), 'UM')
    call ring_close(r)
    UM = e(1)
  end function UM

  ! UMprime (:50-89): the adjoint gradient
  subroutine UMprime(x, answer, a, b)
    double precision, intent(in):: x(:,:,:)
    double precision, intent(out), target:: answer(:,:,:)
    double precision, intent(in), optional:: a(:,:), b(:,:)
    double precision:: e
    call UMforceenergy(x, answer, e, a, b)
  end subroutine UMprime

  ! UMforceenergy (:93-151)
  subroutine UMforceenergy(x, answer, energy, a, b)
    double precision, intent(in):: x(:,:,:)
    double precision, intent(out), target:: answer(:,:,:)
    double precision, intent(out):: energy
    double precision, intent(in), optional:: a(:,:), b(:,:)
    double precision, target:: e(1)
    type(c_ptr):: r
    r = ring_open(a, b)
    call rp_check(rp_ring_set_x(r, x), 'UMforceenergy')
    call rp_check(rp_ring_um(r, c_loc(e), c_loc(answer)), 'UMforceenergy')
    call ring_close(r)
    energy = e(1)
  end subroutine UMforceenergy

  ! UMhessian (:155-217): mass-weighted band AB(ndof+1, n*ndof), the reference's coupling
  subroutine UMhessian(x, singlewell, answer)
    double precision, intent(in):: x(:,:,:)
    logical, intent(in):: singlewell
    double precision, intent(out):: answer(:,:)
    type(c_ptr):: r
    r = ring_open()
    call rp_check(rp_ring_set_x(r, x), 'UMhessian')
    call rp_check(rp_ring_hessian_band(r, 0_c_int, merge(1_c_int, 0_c_int, singlewell), &
         answer), 'UMhessian')
    call ring_close(r)
  end subroutine UMhessian

  ! instanton (:679-777): m = 8, factr = 1e6, pgtol = eps2, unbounded, no iteration cap
  subroutine instanton(xtilde, a, b)
    double precision, intent(inout):: xtilde(:,:,:)
    double precision, intent(in), optional:: a(:,:), b(:,:)
    double precision:: act(1)
    integer(c_int):: iters(1), status(1)
    type(c_ptr):: r
    r = ring_open(a, b)
    call rp_check(rp_instanton(r, eps2, 1.0d6, 8_c_int, 0_c_int, xtilde, act, iters, status), &
         'instanton')
    call ring_close(r)
    if (status(1) /= 0) then
       write(*,*) "Error:"
       write(*,*) "instanton: L-BFGS status", status(1), "after", iters(1), "iterations"
    end if
  end subroutine instanton

  ! detJ (:782-827): the ascending spectrum of the band matrix at x, as its callers use it
  ! (see the header); interphess is ignored, as the reference ignores it (:815)
  subroutine detJ(x, etasquared, singlewell, interphess, eigvecs)
    double precision, intent(in):: x(:,:,:)
    double precision, intent(inout):: etasquared(:)
    logical, intent(in):: singlewell
    double precision, intent(in), optional:: interphess(:,:,:)
    double precision, intent(out), optional:: eigvecs(:,:)
    double precision:: ld(1), e1(1), rest
    double precision, allocatable:: low(:)
    integer(c_int):: nn(1), k
    integer:: i, m
    type(c_ptr):: r
    if (present(eigvecs)) then
       write(*,*) "detJ: eigenvectors are not formed on this path"
       stop 1
    end if
    m = size(etasquared)
    r = ring_open()
    call rp_check(rp_ring_set_x(r, x), 'detJ')
    call rp_check(rp_ring_logdet(r, merge(1_c_int, 0_c_int, singlewell), 0_c_int, ld, nn, e1), 'detJ')
    k = int(min(nn(1) + 1, m), c_int)
    allocate(low(k))
    call rp_check(rp_ring_lowest_eigs(r, merge(1_c_int, 0_c_int, singlewell), k, low), 'detJ')
    call ring_close(r)
    etasquared(1:k) = low
    if (m > k) then
       ! ld = sum of ln(eta) over every eta > 0; the k lowest carry their own share
       rest = ld(1)
       do i = 1, k
          if (low(i) > 0.0d0) rest = rest - log(low(i))
       end do
       etasquared(k+1:m) = exp(rest/dble(m - k))
    end if
  end subroutine detJ

  ! -------------------------------------------------------------------- host helpers
  ! The rotation rotate_atoms applies about axis ax (:317-376): with (j, k) the other two
  ! axes in increasing order, v_j <- cos(t) v_j + sin(t) v_k, v_k <- cos(t) v_k - sin(t) v_j.
  subroutine rot_pair(v, ax, t)
    double precision, intent(inout):: v(3)
    integer, intent(in):: ax
    double precision, intent(in):: t
    integer:: j, k
    double precision:: c, s, vj, vk
    select case (ax)
    case (1)
       j = 2; k = 3
    case (2)
       j = 1; k = 3
    case default
       j = 1; k = 2
    end select
    c = cos(t)
    s = sin(t)
    vj = v(j)
    vk = v(k)
    v(j) = c*vj + s*vk
    v(k) = c*vk - s*vj
  end subroutine rot_pair

  subroutine rotate_vec(vec, axis, theta)
    double precision:: vec(:), theta
    integer:: axis
    call rot_pair(vec(1:3), axis, theta)
  end subroutine rotate_vec

  ! every atom turned about axis by theta; angles below 1e-10 leave the atoms as they are
  subroutine rotate_atoms(atoms, axis, theta)
    double precision, intent(in):: theta
    integer, intent(in):: axis
    double precision, intent(inout):: atoms(:,:)
    integer:: i
    if (abs(theta) <= 1.0d-10) return
    do i = 1, natom
       call rot_pair(atoms(1:3, i), axis, theta)
    end do
  end subroutine rotate_atoms

  ! get_align (:222-265): the angles that put atom1 at the origin, atom2 on +x and atom3 in
  ! the xz half-plane z >= 0 (rotations about z, then y, then x)
  subroutine get_align(atomsin, theta1, theta2, theta3, origin)
    double precision, intent(in):: atomsin(:,:)
    double precision, intent(out):: theta1, theta2, theta3, origin(ndim)
    double precision:: atoms(ndim, natom), w(3)
    integer:: i
    if (ndim /= 3) then
       write(*,*) "Wrong number of dimensions; change align_atoms subroutine!"
       stop
    end if
    origin(:) = atomsin(:, atom1)
    do i = 1, natom
       atoms(:, i) = atomsin(:, i) - origin(:)
    end do
    w = atoms(:, atom2) - atoms(:, atom1)
    theta1 = atan2(w(2), w(1))
    call rotate_atoms(atoms, 3, theta1)
    w = atoms(:, atom2) - atoms(:, atom1)
    theta2 = atan2(w(3), w(1))
    call rotate_atoms(atoms, 2, theta2)
    w = atoms(:, atom3) - atoms(:, atom1)
    theta3 = -atan2(w(2), w(3))
  end subroutine get_align

  ! align_atoms (:269-312): move atomsin's own atom1 to the origin and apply the three
  ! rotations; when every angle is below 1e-10 the atoms are copied unmoved
  subroutine align_atoms(atomsin, theta1, theta2, theta3, origin, atomsout)
    double precision, intent(in):: atomsin(:,:), theta1, theta2, theta3, origin(ndim)
    double precision, intent(out):: atomsout(:,:)
    integer:: i
    if (ndim /= 3) then
       write(*,*) "Wrong number of dimensions; change align_atoms subroutine!"
       stop
    end if
    if (abs(theta1) > 1.0d-10 .or. abs(theta2) > 1.0d-10 .or. abs(theta3) > 1.0d-10) then
       do i = 1, natom
          atomsout(:, i) = atomsin(:, i) - atomsin(:, atom1)
       end do
       call rotate_atoms(atomsout, 3, theta1)
       call rotate_atoms(atomsout, 2, theta2)
       call rotate_atoms(atomsout, 1, theta3)
    else
       atomsout(:,:) = atomsin(:,:)
    end if
  end subroutine align_atoms

  ! Cubic spline second derivatives (Numerical Recipes' spline, :397-442): natural ends when
  ! yp1 / ypn >= 0.99e30, else the given end slopes; the tridiagonal system by elimination.
  subroutine spline(x, y, yp1, ypn, y2)
    double precision, dimension(:), intent(in):: x, y
    double precision, intent(in):: yp1, ypn
    double precision, dimension(:), intent(out):: y2
    double precision:: lo(size(x)), di(size(x)), up(size(x)), rhs(size(x)), piv
    integer:: m, i
    m = size(x)
    if (size(y) /= m .or. size(y2) /= m) then
       write(*,*) 'nrerror: an assert_eq failed with this tag:', 'spline'
       stop 'program terminated by assert_eq'
    end if
    ! rows i = 2..m-1: h_{i-1} y2_{i-1} + 2 (h_{i-1} + h_i) y2_i + h_i y2_{i+1}
    !                   = 6 [(y_{i+1} - y_i)/h_i - (y_i - y_{i-1})/h_{i-1}]
    do i = 2, m - 1
       lo(i) = x(i) - x(i-1)
       up(i) = x(i+1) - x(i)
       di(i) = 2.0d0*(lo(i) + up(i))
       rhs(i) = 6.0d0*((y(i+1) - y(i))/up(i) - (y(i) - y(i-1))/lo(i))
    end do
    di(1) = 1.0d0
    di(m) = 1.0d0
    lo(1) = 0.0d0
    up(m) = 0.0d0
    if (yp1 > 0.99d30) then
       up(1) = 0.0d0
       rhs(1) = 0.0d0
    else
       up(1) = 0.5d0
       rhs(1) = (3.0d0/(x(2) - x(1)))*((y(2) - y(1))/(x(2) - x(1)) - yp1)
    end if
    if (ypn > 0.99d30) then
       lo(m) = 0.0d0
       rhs(m) = 0.0d0
    else
       lo(m) = 0.5d0
       rhs(m) = (-3.0d0/(x(m) - x(m-1)))*((y(m) - y(m-1))/(x(m) - x(m-1)) - ypn)
    end if
    ! forward elimination, then back substitution
    do i = 2, m
       piv = lo(i)/di(i-1)
       di(i) = di(i) - piv*up(i-1)
       rhs(i) = rhs(i) - piv*rhs(i-1)
    end do
    y2(m) = rhs(m)/di(m)
    do i = m - 1, 1, -1
       y2(i) = (rhs(i) - up(i)*y2(i+1))/di(i)
    end do
  end subroutine spline

  ! locate (:571-603): j with xx(j) <= x < xx(j+1) by bisection (either table order);
  ! 0 or size(xx) when x is outside, 1 and n-1 at the end points
  integer function locate(xx, x)
    double precision, dimension(:), intent(in):: xx
    double precision, intent(in):: x
    integer:: m, jl, ju, jm
    logical:: ascnd
    m = size(xx)
    ascnd = xx(m) >= xx(1)
    jl = 0
    ju = m + 1
    do while (ju - jl > 1)
       jm = (ju + jl)/2
       if (ascnd .eqv. (x >= xx(jm))) then
          jl = jm
       else
          ju = jm
       end if
    end do
    if (x == xx(1)) then
       locate = 1
    else if (x == xx(m)) then
       locate = m - 1
    else
       locate = jl
    end if
  end function locate

  ! the interval of x and its local coordinates (shared by splint and splin_grad)
  subroutine spline_cell(xa, ya, y2a, x, klo, aa, bb, h, tag)
    double precision, dimension(:), intent(in):: xa, ya, y2a
    double precision, intent(in):: x
    integer, intent(out):: klo
    double precision, intent(out):: aa, bb, h
    character(len=*), intent(in):: tag
    integer:: m
    m = size(xa)
    if (size(ya) /= m .or. size(y2a) /= m) then
       write(*,*) 'nrerror: an assert_eq failed with this tag:', tag
       stop 'program terminated by assert_eq'
    end if
    klo = max(min(locate(xa, x), m - 1), 1)
    h = xa(klo+1) - xa(klo)
    if (h == 0.0d0) then
       write(*,*) "bad xa input in ", tag
       stop
    end if
    aa = (xa(klo+1) - x)/h
    bb = (x - xa(klo))/h
  end subroutine spline_cell

  ! splint (:498-528): the cubic spline's value at x
  double precision function splint(xa, ya, y2a, x)
    double precision, dimension(:), intent(in):: xa, ya, y2a
    double precision, intent(in):: x
    integer:: k
    double precision:: aa, bb, h
    call spline_cell(xa, ya, y2a, x, k, aa, bb, h, 'splint')
    splint = aa*ya(k) + bb*ya(k+1) + ((aa**3 - aa)*y2a(k) + (bb**3 - bb)*y2a(k+1))*(h**2)/6.0d0
  end function splint

  ! splin_grad (:532-561): the cubic spline's slope at x
  double precision function splin_grad(xa, ya, y2a, x)
    double precision, dimension(:), intent(in):: xa, ya, y2a
    double precision, intent(in):: x
    integer:: k
    double precision:: aa, bb, h
    call spline_cell(xa, ya, y2a, x, k, aa, bb, h, 'splin_grad')
    splin_grad = (ya(k+1) - ya(k))/h + ((1.0d0 - 3.0d0*aa**2)*y2a(k) + (3.0d0*bb**2 - 1.0d0)*y2a(k+1))*h/6.0d0
  end function splin_grad

  ! findmiddle (:832-871): where the spline of path(lampath) has zero slope in [x1, x2], by
  ! bisection to 1e-6 in at most 40 halvings
  double precision function findmiddle(x1, x2, lampath, path, npath)
    integer:: npath
    double precision:: x1, x2, lampath(:), path(:)
    double precision:: y2(npath), f1, f2, dx, xm, fm
    integer:: it
    call spline(lampath, path, 1.0d31, 1.0d31, y2)
    f2 = splin_grad(lampath, path, y2, x2)
    f1 = splin_grad(lampath, path, y2, x1)
    if (f1*f2 >= 0.0d0) then
       write(*,*) 'root must be bracketed in findmiddle'
       write(*,*) 'you probably did a terrible job at finding the right path'
       write(*,*) x1, f1
       write(*,*) x2, f2
       stop
    end if
    ! start from the end with the negative slope and step towards the other
    if (f1 < 0.0d0) then
       findmiddle = x1
       dx = x2 - x1
    else
       findmiddle = x2
       dx = x1 - x2
    end if
    do it = 1, 40
       dx = 0.5d0*dx
       xm = findmiddle + dx
       fm = splin_grad(lampath, path, y2, xm)
       if (fm <= 0.0d0) findmiddle = xm
       if (abs(dx) < 1.0d-6 .or. fm == 0.0d0) return
    end do
    write(*,*) 'too many bisections in findmiddle'
    stop
  end function findmiddle

  ! eucliddist (:876-890)
  double precision function eucliddist(x1, x2)
    double precision, intent(in):: x1(:,:), x2(:,:)
    eucliddist = sqrt(sum((x1(1:ndim, 1:natom) - x2(1:ndim, 1:natom))**2))
  end function eucliddist

  ! read_path (:895-1035): the frames of path.xyz (units by xunit), aligned with the first
  ! frame's angles; lampath = normalised arc length, Vpath = V per frame; xtilde = the path
  ! spline at n evenly spaced points.  instapath: the instanton from xtilde becomes the path
  ! (with the wells as its ends when fixedends).  centre: lampath remapped so the slope's
  ! zero sits at 1/2.  splinepath: the final spline.
  subroutine read_path(instapath, centre, npath, path, Vpath, splinepath, lampath)
    integer, intent(inout):: npath
    logical, intent(in):: instapath, centre
    double precision, allocatable, intent(out):: path(:,:,:), Vpath(:), splinepath(:,:,:), lampath(:)
    double precision, allocatable:: frame(:,:)
    double precision:: origin(3), t1, t2, t3, xm, qa, qb
    integer:: i, j, k, u, idummy
    character:: clabel
    character(len=256):: line
    allocate(lampath(npath), frame(ndim, natom))
    allocate(path(npath, ndim, natom), Vpath(npath), splinepath(npath, ndim, natom))
    path = 0.0d0
    t1 = 0.0d0; t2 = 0.0d0; t3 = 0.0d0
    origin = 0.0d0
    open(newunit=u, file="path.xyz", status='old', action='read')
    do i = 1, npath
       read(u,*) idummy
       read(u,'(A)') line
       do j = 1, natom
          read(u,*) clabel, (frame(k, j), k = 1, ndim)
       end do
       if (xunit == 2) frame = frame/0.529177d0
       if (i == 1) call get_align(frame, t1, t2, t3, origin)
       call align_atoms(frame, t1, t2, t3, origin, path(i,:,:))
       if (i == 1) then
          lampath(1) = 0.0d0
       else
          lampath(i) = lampath(i-1) + eucliddist(path(i-1,:,:), path(i,:,:))
       end if
       Vpath(i) = V(path(i,:,:))
    end do
    close(u)
    lampath = lampath/lampath(npath)
    do i = 1, ndim
       do j = 1, natom
          call spline(lampath, path(:,i,j), 1.0d31, 1.0d31, splinepath(:,i,j))
          do k = 1, n
             xtilde(k,i,j) = splint(lampath, path(:,i,j), splinepath(:,i,j), dble(k-1)/dble(n-1))
          end do
       end do
    end do
    if (instapath) then
       write(*,*) "Finding instanton"
       if (fixedends) then
          call instanton(xtilde, well1, well2)
       else
          call instanton(xtilde)
       end if
       deallocate(path, lampath, Vpath, splinepath)
       if (fixedends) then
          npath = n + 2
       else
          npath = n
       end if
       allocate(path(npath, ndim, natom), Vpath(npath), splinepath(npath, ndim, natom), lampath(npath))
       if (fixedends) then
          path(2:n+1,:,:) = xtilde
          path(1,:,:) = well1
          path(npath,:,:) = well2
       else
          path = xtilde
       end if
       write(*,*) "Found instanton."
       open(newunit=u, file="instanton.xyz")
       do i = 1, n
          write(u,*) natom
          write(u,*) "Energy of minimum", i
          do j = 1, natom
             write(u,*) label(j), (xtilde(i,k,j)*0.529177d0, k = 1, ndim)
          end do
       end do
       close(u)
       lampath(1) = 0.0d0
       Vpath(1) = V(path(1,:,:))
       do i = 2, npath
          lampath(i) = lampath(i-1) + eucliddist(path(i-1,:,:), path(i,:,:))
          Vpath(i) = V(path(i,:,:))
       end do
       lampath = lampath/lampath(npath)
    end if
    if (centre) then
       ! lampath -> the root of a (l')^2 + b l' = l with l'(xm) = 1/2 (a + b = 1)
       xm = findmiddle(0.3d0, 0.7d0, lampath, Vpath, npath)
       qa = 2.0d0 - 4.0d0*xm
       qb = 4.0d0*xm - 1.0d0
       do i = 1, npath
          if (qa >= 0.0d0) then
             lampath(i) = -0.5d0*qb/qa + sqrt(lampath(i)/qa + (0.5d0*qb/qa)**2)
          else
             lampath(i) = -0.5d0*qb/qa - sqrt(lampath(i)/qa + (0.5d0*qb/qa)**2)
          end if
       end do
       write(*,*) "Centred,", qa, qb, xm
    end if
    do i = 1, ndim
       do j = 1, natom
          call spline(lampath, path(:,i,j), 1.0d31, 1.0d31, splinepath(:,i,j))
       end do
    end do
    open(newunit=u, file="aligned.xyz")
    do i = 1, n
       write(u,*) natom
       write(u,*) "rotation angles:", t1, t2, t3
       do j = 1, natom
          write(u,*) label(j), (xtilde(i,k,j)*0.529177d0, k = 1, ndim)
       end do
    end do
    close(u)
    deallocate(frame)
  end subroutine read_path

  ! read_hess (instantonmod.f90:1039-1074) splines Hessians along the path for the readhess
  ! initial momenta, which this build does not support: a caller that asks for them stops
  ! here, as rptunnel's deck reader rejects readhess.  The routine exists so the reference's
  ! pimd_par.f90 (:202-205) links against this module.
  subroutine read_hess(npath, path, lampath, splinehess, hesspath)
    integer, intent(in):: npath
    double precision, allocatable, intent(in):: path(:,:,:), lampath(:)
    double precision, allocatable, intent(out):: splinehess(:,:,:), hesspath(:,:,:)
    write(*,*) "read_hess: readhess is not supported by the device build (npath =", npath, ")"
    stop
  end subroutine read_hess

  ! The lndetJ sums of rpi_ser.f90 at chain x without forming the spectrum: skip_lowest =
  ! .false. gives lndetJ0's sum over every eta > 0 (:227-236), .true. lndetJ's sum over
  ! eta_2.. > 0 (:355-365); nneg = the inertia's negative count, eta1 the lowest eigenvalue.
  subroutine detJ_logsum(x, singlewell, skip_lowest, logdet, nneg, eta1)
    double precision, intent(in):: x(:,:,:)
    logical, intent(in):: singlewell, skip_lowest
    double precision, intent(out):: logdet, eta1
    integer, intent(out):: nneg
    double precision:: ld(1), e1(1)
    integer(c_int):: nn(1)
    type(c_ptr):: r
    r = ring_open()
    call rp_check(rp_ring_set_x(r, x), 'detJ_logsum')
    call rp_check(rp_ring_logdet(r, merge(1_c_int, 0_c_int, singlewell), &
         merge(1_c_int, 0_c_int, skip_lowest), ld, nn, e1), 'detJ_logsum')
    call ring_close(r)
    logdet = ld(1)
    nneg = nn(1)
    eta1 = e1(1)
  end subroutine detJ_logsum

end module instantonmod
