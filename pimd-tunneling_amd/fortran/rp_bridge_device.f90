! rp_bridge_device.f90 -- `module rp_bridge` for the device plugin (mcmod_rp.f90), whose
! potential the library evaluates in its own kernels.  The drop-in verletint and
! instantonmod reach the plugin only through this module, so the same two modules serve
! a device plugin (this file) or any host plugin of the reference's API (rp_bridge_host.f90).
module rp_bridge
  use iso_c_binding
  use rp_hip
  use mcmod_mass
  implicit none
  private
  public :: rp_bridge_pes, rp_bridge_sync_v0, rp_bridge_restore_v0

contains

  function rp_bridge_pes() result(h)
    type(c_ptr) :: h
    h = rp_pes_handle
  end function rp_bridge_pes

  ! the driver's V0 goes to the device while a chain routine that reports energies runs
  subroutine rp_bridge_sync_v0()
    call rp_sync_v0()
  end subroutine rp_bridge_sync_v0

  subroutine rp_bridge_restore_v0()
    call rp_restore_v0()
  end subroutine rp_bridge_restore_v0

end module rp_bridge
