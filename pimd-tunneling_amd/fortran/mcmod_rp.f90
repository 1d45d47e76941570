! mcmod_rp.f90 -- a `module mcmod_mass` PES plugin whose potential lives on the GPU.
!
! Same public API as the reference's plugins (mcmod_waterdimer.f90:1-105): the module
! variables V0, eps2, atom1..3, label, n, ndim, ndof, natom, xunit, totdof,
! potforcepresent, basename and the procedures V_init, V, Vprime, potforce, Vdoubleprime.
! Each procedure evaluates one bead through libpimdtunnel.so; the batched callers
! (verletint and instantonmod in this directory) hand the device the whole chain
! through rp_pes_handle instead.
!
! V_init reads the potential from the file rp_pes.dat unless the caller has already
! called rp_plugin_setup:
!     kind ndim natom ndp nip
!     dp(1:ndp)
!     ip(1:nip)                        (only when nip > 0)
! Kinds and parameter layouts: RP_PES_* in rp_hip.f90, described in csrc/rp_pes.h.
module mcmod_mass
  use iso_c_binding
  use rp_hip
  implicit none
  double precision::               V0, eps2=1.0d-5
  integer::                        atom1, atom2, atom3
  character, allocatable::         label(:)
  integer::                        n, ndim, ndof, natom, xunit, totdof
  logical::                        potforcepresent
  character(len=20)::              basename
  type(c_ptr)::                    rp_pes_handle = c_null_ptr

contains

  ! Create the device potential from arrays (instead of rp_pes.dat).
  subroutine rp_plugin_setup(kind, ndim_in, natom_in, dp, ip)
    integer, intent(in)::          kind, ndim_in, natom_in
    double precision, intent(in):: dp(:)
    integer, intent(in)::          ip(:)
    integer(c_int), allocatable::  ipc(:)
    if (c_associated(rp_pes_handle)) call rp_check(rp_pes_destroy(rp_pes_handle), 'rp_plugin_setup')
    call rp_check(rp_init(0_c_int), 'rp_plugin_setup')
    allocate(ipc(max(size(ip), 1)))
    ipc = 0
    if (size(ip) > 0) ipc(1:size(ip)) = int(ip, c_int)
    call rp_check(rp_pes_create(int(kind, c_int), int(ndim_in, c_int), int(natom_in, c_int), &
         dp, int(size(dp), c_int), ipc, int(size(ip), c_int), rp_pes_handle), 'rp_plugin_setup')
    ! the plugin subtracts V0 itself, so the device potential keeps V0 = 0 except while a
    ! chain routine runs (see rp_sync_v0)
    call rp_check(rp_pes_set_v0(rp_pes_handle, 0.0d0), 'rp_plugin_setup')
  end subroutine rp_plugin_setup

  subroutine V_init(iproc)
    integer, intent(in):: iproc
    integer:: kind, nd, na, ndp, nip, u
    double precision, allocatable:: dp(:)
    integer, allocatable:: ip(:)
    potforcepresent = .true.
    V0 = 0.0d0
    if (c_associated(rp_pes_handle)) return
    open(newunit=u, file='rp_pes.dat', status='old', action='read')
    read(u,*) kind, nd, na, ndp, nip
    allocate(dp(ndp), ip(nip))
    read(u,*) dp
    if (nip > 0) read(u,*) ip
    close(u)
    call rp_plugin_setup(kind, nd, na, dp, ip)
    if (iproc == 0) write(*,*) 'mcmod_rp: device potential kind', kind, 'with', nd*na, 'dof'
  end subroutine V_init

  ! Pass the driver's V0 to the device before a chain routine that reports energies.
  subroutine rp_sync_v0()
    call rp_check(rp_pes_set_v0(rp_pes_handle, V0), 'rp_sync_v0')
  end subroutine rp_sync_v0

  subroutine rp_restore_v0()
    call rp_check(rp_pes_set_v0(rp_pes_handle, 0.0d0), 'rp_restore_v0')
  end subroutine rp_restore_v0

  function V(x)
    double precision:: V, x(:,:)
    double precision:: e(1), g(size(x))
    call rp_check(rp_pes_potforce(rp_pes_handle, 1_c_int, x, g, e), 'V')
    V = e(1) - V0
  end function V

  subroutine Vprime(x, grad)
    double precision:: grad(:,:), x(:,:)
    double precision:: e(1), g(size(x))
    call rp_check(rp_pes_potforce(rp_pes_handle, 1_c_int, x, g, e), 'Vprime')
    grad = reshape(g, shape(grad))
  end subroutine Vprime

  subroutine potforce(x, grad, energy)
    double precision:: grad(:,:), x(:,:), energy
    double precision:: e(1), g(size(x))
    call rp_check(rp_pes_potforce(rp_pes_handle, 1_c_int, x, g, e), 'potforce')
    grad = reshape(g, shape(grad))
    energy = e(1) - V0
  end subroutine potforce

  ! hess(ndim, natom, ndim, natom) = d grad(k) / d x(i): exact second derivatives from the
  ! device's forward-over-reverse, where the reference's plugins difference Vprime
  subroutine Vdoubleprime(x, hess)
    double precision:: hess(:,:,:,:), x(:,:)
    double precision:: h(size(x), size(x))
    call rp_check(rp_pes_hessian(rp_pes_handle, 1_c_int, x, h), 'Vdoubleprime')
    ! device layout hess[i][k] (C row-major) = Fortran h(k, i); the matrix is symmetric
    hess = reshape(transpose(h), shape(hess))
  end subroutine Vdoubleprime

end module mcmod_mass
