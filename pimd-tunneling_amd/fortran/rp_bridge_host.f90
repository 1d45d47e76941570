! rp_bridge_host.f90 -- `module rp_bridge` for a plugin that lives on the host.
!
! Link this with ANY `module mcmod_mass` written to the reference's plugin API
! (mcmod_waterdimer.f90:1-105: V, Vprime, potforce, Vdoubleprime, potforcepresent) and the
! drop-in verletint / instantonmod run their device kernels around that plugin: the library
! calls back here with whole chains x(nbead, ndim, natom), and these callbacks loop over the
! beads the way the reference's step_v (verletmodule.f90:573) and UMforceenergy
! (instantonmod.f90:105) do.  The plugin subtracts its own V0, so nothing is synchronised.
! The alternative, rp_bridge_device.f90, is for the device plugin mcmod_rp.f90; a program
! links exactly one of the two, as it links exactly one plugin (makefile:19-80).
module rp_bridge
  use iso_c_binding
  use rp_hip
  use mcmod_mass
  implicit none
  private
  public :: rp_bridge_pes, rp_bridge_sync_v0, rp_bridge_restore_v0
  type(c_ptr), save :: handle = c_null_ptr

contains

  ! the library's handle of the plugin, created on first use
  function rp_bridge_pes() result(h)
    type(c_ptr) :: h
    if (.not. c_associated(handle)) then
       call rp_check(rp_init(0_c_int), 'rp_bridge')
       call rp_check(rp_pes_create_host(int(ndim, c_int), int(natom, c_int), &
            c_funloc(bridge_potforce), c_funloc(bridge_hessian), c_null_ptr, handle), 'rp_bridge')
    end if
    h = handle
  end function rp_bridge_pes

  subroutine rp_bridge_sync_v0()
  end subroutine rp_bridge_sync_v0

  subroutine rp_bridge_restore_v0()
  end subroutine rp_bridge_restore_v0

  ! energies and gradients of nbead beads: potforce when the plugin has it, else V and Vprime
  ! (instantonmod.f90:103-111)
  integer(c_int) function bridge_potforce(user, nbead, x, grad, energy) bind(C)
    type(c_ptr), value :: user
    integer(c_int), value :: nbead
    real(c_double), intent(in) :: x(nbead, ndim, natom)
    real(c_double), intent(out) :: grad(nbead, ndim, natom), energy(nbead)
    double precision :: xb(ndim, natom), gb(ndim, natom)
    integer :: b
    do b = 1, nbead
       xb = x(b, :, :)
       if (potforcepresent) then
          call potforce(xb, gb, energy(b))
       else
          energy(b) = V(xb)
          call Vprime(xb, gb)
       end if
       grad(b, :, :) = gb
    end do
    bridge_potforce = 0
  end function bridge_potforce

  ! Vdoubleprime per bead: hess(i,j,k,l) = d grad(k,l) / d x(i,j), returned as the C array
  ! out[bead][p][q] = d grad_q / d x_p, i.e. out(q, p, bead) here
  integer(c_int) function bridge_hessian(user, nbead, x, out) bind(C)
    type(c_ptr), value :: user
    integer(c_int), value :: nbead
    real(c_double), intent(in) :: x(nbead, ndim, natom)
    real(c_double), intent(out) :: out(ndim*natom, ndim*natom, nbead)
    double precision :: xb(ndim, natom), hb(ndim, natom, ndim, natom)
    integer :: b
    do b = 1, nbead
       xb = x(b, :, :)   ! a copy: differencing plugins perturb x in place
       call Vdoubleprime(xb, hb)
       out(:, :, b) = transpose(reshape(hb, (/ndim*natom, ndim*natom/)))
    end do
    bridge_hessian = 0
  end function bridge_hessian

end module rp_bridge
