! instantonmod_rp.f90 -- `module instantonmod` with the ring-polymer instanton routines
! of the reference (instantonmod.f90:1-217, 679-827) evaluated on the GPU.
!
! Same module variables (beta, betan, UMtilde, xtilde, well1, well2, mass, fixedends) and
! the same calls: UM, UMprime, UMforceenergy, UMhessian, instanton.  detJ's role (the
! eigenvalues feeding the lndetJ sums of rpi_ser.f90:221-237, 354-365) is taken by
! detJ_logsum, which returns those sums directly from a block LDL^T factorisation of the
! band matrix with inertia counting, so no eigen-decomposition is formed.  Chains use the
! reference layout x(n, ndim, natom); n, ndim, natom come from mcmod_mass.
module instantonmod
  use iso_c_binding
  use rp_hip
  use mcmod_mass
  implicit none
  double precision, parameter::    pi=3.14159265358979d0
  double precision::               beta, betan, UMtilde
  double precision, allocatable::  xtilde(:,:,:), well1(:,:), well2(:,:), mass(:)
  logical::                        fixedends

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
    call rp_sync_v0()
    call rp_check(rp_ring_create(rp_pes_handle, mass, 1_c_int, nn, bn, aa, bb, &
         merge(1_c_int, 0_c_int, fixedends), r), 'instantonmod')
  end function ring_open

  subroutine ring_close(r)
    type(c_ptr), intent(inout):: r
    call rp_check(rp_ring_destroy(r), 'instantonmod')
    r = c_null_ptr
    call rp_restore_v0()
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

  ! instanton (:679-777): m = 8, factr = 1e6, pgtol = eps2, unbounded
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

  ! The lndetJ sums of rpi_ser.f90 at chain x: skip_lowest = .false. gives lndetJ0's sum
  ! over every eta > 0 (:227-236), .true. lndetJ's sum over eta_2.. > 0 (:355-365).
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
