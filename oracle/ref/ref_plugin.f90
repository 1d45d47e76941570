! ref_plugin.f90 -- a `module mcmod_mass` plugin with the config potentials, written for
! this project against the plugin API of the reference (mcmod_waterdimer.f90:1-105), so
! that the reference's own instantonmod.f90 can be compiled here and run as a
! cross-check of the oracle.  TEST INFRASTRUCTURE ONLY (built into oracle/_ref/).
!
! Potentials (parameter layouts as in oracle/oracle.c):
!   kind 1 quartic   dp = (h, x0)                 V = sum h((x/x0)^2-1)^2
!   kind 2 coupled   dp = (V0, a, m, w, c)        V = V0(x^2/a^2-1)^2 + m w^2 (y+c x^2)^2/2
!   kind 3 EVB       dp = (kappa, Q(D,D), lam(D), cub(D), xa(D)), ip = perm(natom)
!   kind 5 harmonic  dp = (k(D), c(D))
! Vdoubleprime is analytic for kinds 1, 2, 5 and central differences of Vprime with
! eps = 1e-4 for the EVB, as the reference's plugins do (mcmod_waterdimer.f90:83-103).
module mcmod_mass
  implicit none
  double precision::               V0, eps2=1.0d-5
  integer::                        atom1, atom2, atom3
  character, allocatable::         label(:)
  integer::                        n, ndim, ndof, natom, xunit, totdof
  logical::                        potforcepresent
  character(len=20)::              basename
  integer::                        pes_kind = 0
  double precision, allocatable::  pes_dp(:)
  integer, allocatable::           pes_ip(:)

contains

  ! A driver that sets no potential (the reference's own drivers only call V_init) reads it
  ! from rp_pes.dat, in the format of the device plugin mcmod_rp.f90:
  !     kind ndim natom ndp nip / dp(1:ndp) / ip(1:nip), ip 0-based atom indices
  subroutine V_init(iproc)
    integer, intent(in):: iproc
    integer:: nd, na, ndp, nip, u
    potforcepresent=.true.
    V0=0.0d0
    if (pes_kind /= 0) return
    open(newunit=u, file='rp_pes.dat', status='old', action='read')
    read(u,*) pes_kind, nd, na, ndp, nip
    allocate(pes_dp(ndp), pes_ip(max(nip,1)))
    read(u,*) pes_dp
    if (nip > 0) read(u,*) pes_ip(1:nip)
    close(u)
  end subroutine V_init

  ! energy and gradient in dof order, before the V0 shift
  subroutine eval_pes(xv, e, gv)
    double precision, intent(in)::  xv(:)
    double precision, intent(out):: e, gv(:)
    double precision:: t, s, va, vb, dl, r
    double precision, allocatable:: px(:), ga(:), gb(:), gpx(:)
    integer:: d, a, dd
    select case (pes_kind)
    case (1)
       e = 0.0d0
       do d = 1, ndof
          t = (xv(d)/pes_dp(2))**2 - 1.0d0
          e = e + pes_dp(1)*t*t
          gv(d) = t*4.0d0*pes_dp(1)*xv(d)/pes_dp(2)**2
       end do
    case (2)
       t = xv(1)**2/pes_dp(2)**2 - 1.0d0
       s = xv(2) + pes_dp(5)*xv(1)**2
       e = pes_dp(1)*t*t + 0.5d0*pes_dp(3)*pes_dp(4)**2*s*s
       gv(1) = 4.0d0*pes_dp(1)*t*xv(1)/pes_dp(2)**2 + pes_dp(3)*pes_dp(4)**2*s*2.0d0*pes_dp(5)*xv(1)
       gv(2) = pes_dp(3)*pes_dp(4)**2*s
    case (3)
       allocate(px(ndof), ga(ndof), gb(ndof), gpx(ndof))
       do a = 1, natom
          do d = 1, ndim
             px((a-1)*ndim + d) = xv(pes_ip(a)*ndim + d)
          end do
       end do
       call evb_well(xv, va, ga)
       call evb_well(px, vb, gpx)
       do a = 1, natom
          do d = 1, ndim
             gb(pes_ip(a)*ndim + d) = gpx((a-1)*ndim + d)
          end do
       end do
       dl = va - vb
       r = sqrt(0.25d0*dl*dl + pes_dp(1)**2)
       e = 0.5d0*(va + vb) - r
       do dd = 1, ndof
          gv(dd) = 0.5d0*(ga(dd) + gb(dd)) - dl*(ga(dd) - gb(dd))/(4.0d0*r)
       end do
       deallocate(px, ga, gb, gpx)
    case (5)
       e = 0.0d0
       do d = 1, ndof
          e = e + 0.5d0*pes_dp(d)*(xv(d) - pes_dp(ndof + d))**2
          gv(d) = pes_dp(d)*(xv(d) - pes_dp(ndof + d))
       end do
    case default
       write(*,*) "ref_plugin: unknown potential", pes_kind
       stop
    end select
  end subroutine eval_pes

  ! VA = y.Lam.y/2 + sum c y^3, y = Q^T (z - xa), and its gradient Q (Lam y + 3 c y^2)
  subroutine evb_well(z, vw, gz)
    double precision, intent(in)::  z(:)
    double precision, intent(out):: vw, gz(:)
    integer:: i, j, qoff, loff, coff, aoff
    double precision:: y, h
    qoff = 1
    loff = 1 + ndof*ndof
    coff = loff + ndof
    aoff = coff + ndof
    vw = 0.0d0
    gz = 0.0d0
    do i = 1, ndof
       y = 0.0d0
       do j = 1, ndof
          y = y + pes_dp(qoff + j + ndof*(i-1))*(z(j) - pes_dp(aoff + j))
       end do
       vw = vw + 0.5d0*pes_dp(loff + i)*y*y + pes_dp(coff + i)*y*y*y
       h = pes_dp(loff + i)*y + 3.0d0*pes_dp(coff + i)*y*y
       do j = 1, ndof
          gz(j) = gz(j) + pes_dp(qoff + j + ndof*(i-1))*h
       end do
    end do
  end subroutine evb_well

  function V(x)
    double precision:: V, x(:,:)
    double precision, allocatable:: xv(:), gv(:)
    allocate(xv(ndof), gv(ndof))
    xv = reshape(x, (/ndof/))
    call eval_pes(xv, V, gv)
    V = V - V0
    deallocate(xv, gv)
  end function V

  subroutine Vprime(x, grad)
    double precision:: grad(:,:), x(:,:), e
    double precision, allocatable:: xv(:), gv(:)
    allocate(xv(ndof), gv(ndof))
    xv = reshape(x, (/ndof/))
    call eval_pes(xv, e, gv)
    grad = reshape(gv, (/ndim, natom/))
    deallocate(xv, gv)
  end subroutine Vprime

  subroutine potforce(x, grad, energy)
    double precision:: grad(:,:), x(:,:), energy
    double precision, allocatable:: xv(:), gv(:)
    allocate(xv(ndof), gv(ndof))
    xv = reshape(x, (/ndof/))
    call eval_pes(xv, energy, gv)
    energy = energy - V0
    grad = reshape(gv, (/ndim, natom/))
    deallocate(xv, gv)
  end subroutine potforce

  subroutine Vdoubleprime(x, hess)
    double precision:: hess(:,:,:,:), x(:,:), eps, s
    double precision, allocatable:: gradplus(:,:), gradminus(:,:)
    integer:: i, j, d
    hess = 0.0d0
    select case (pes_kind)
    case (1)
       do i = 1, ndim
          do j = 1, natom
             hess(i,j,i,j) = 4.0d0*pes_dp(1)*(3.0d0*x(i,j)**2/pes_dp(2)**2 - 1.0d0)/pes_dp(2)**2
          end do
       end do
    case (2)
       s = x(2,1) + pes_dp(5)*x(1,1)**2
       hess(1,1,1,1) = 4.0d0*pes_dp(1)*(3.0d0*x(1,1)**2/pes_dp(2)**2 - 1.0d0)/pes_dp(2)**2 &
            + pes_dp(3)*pes_dp(4)**2*(4.0d0*pes_dp(5)**2*x(1,1)**2 + 2.0d0*pes_dp(5)*s)
       hess(1,1,2,1) = pes_dp(3)*pes_dp(4)**2*2.0d0*pes_dp(5)*x(1,1)
       hess(2,1,1,1) = hess(1,1,2,1)
       hess(2,1,2,1) = pes_dp(3)*pes_dp(4)**2
    case (5)
       do i = 1, ndim
          do j = 1, natom
             d = (j-1)*ndim + i
             hess(i,j,i,j) = pes_dp(d)
          end do
       end do
    case default
       eps = 1d-4
       allocate(gradplus(ndim, natom), gradminus(ndim, natom))
       do i = 1, ndim
          do j = 1, natom
             x(i,j) = x(i,j) + eps
             call Vprime(x, gradplus)
             x(i,j) = x(i,j) - 2.0d0*eps
             call Vprime(x, gradminus)
             x(i,j) = x(i,j) + eps
             hess(i,j,:,:) = (gradplus(:,:) - gradminus(:,:))/(2.0d0*eps)
          end do
       end do
       deallocate(gradplus, gradminus)
    end select
  end subroutine Vdoubleprime

end module mcmod_mass
