! rp_hip.f90 -- iso_c_binding interfaces of libpimdtunnel.so (include/rp_hip.h).
!
! Array arguments use the reference's own layout, x(n, ndim, natom) column-major, so a
! Fortran caller passes its arrays unchanged.  Handles are type(c_ptr).  rp_check turns
! a non-zero status into the reference's failure behaviour: print the message and stop
! (verletmodule.f90:533-536, 577-580).
module rp_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: RP_OK = 0, RP_ERR_ARG = 1, RP_ERR_HIP = 2, RP_ERR_NAN = 3, &
       RP_ERR_NOISE = 4, RP_ERR_STATE = 5, RP_ERR_CONVERGE = 6
  integer(c_int), parameter :: RP_PES_QUARTIC = 1, RP_PES_COUPLED2D = 2, RP_PES_EVB = 3, &
       RP_PES_WATER = 4, RP_PES_HARMONIC = 5, RP_PES_HOST = 6
  integer(c_int), parameter :: RP_THERMO_ANDERSEN = 1, RP_THERMO_LANGEVIN = 2

  type, bind(C) :: rp_polymer_desc
     integer(c_int)     :: n
     real(c_double)     :: betan
     real(c_double)     :: xi
     type(c_ptr)        :: a
     type(c_ptr)        :: b
     type(c_ptr)        :: dbdl
     integer(c_int64_t) :: stream
  end type rp_polymer_desc

  interface
     ! ---------------------------------------------------------------- runtime
     integer(c_int) function rp_init(device) bind(C, name="rp_init")
       import :: c_int
       integer(c_int), value :: device
     end function rp_init
     integer(c_int) function rp_finalize() bind(C, name="rp_finalize")
       import :: c_int
     end function rp_finalize
     integer(c_int) function rp_last_error(buf, len) bind(C, name="rp_last_error")
       import :: c_int, c_char
       character(kind=c_char) :: buf(*)
       integer(c_int), value :: len
     end function rp_last_error
     integer(c_int) function rp_version() bind(C, name="rp_version")
       import :: c_int
     end function rp_version

     ! ---------------------------------------------------------------- PES plugin
     integer(c_int) function rp_pes_create(kind, ndim, natom, dp, ndp, ip, nip, out) &
          bind(C, name="rp_pes_create")
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: kind, ndim, natom, ndp, nip
       real(c_double), intent(in) :: dp(*)
       integer(c_int), intent(in) :: ip(*)
       type(c_ptr), intent(out) :: out
     end function rp_pes_create
     ! a plugin that lives on the host: potforce / hessian are bind(C) callbacks
     ! (see rp_bridge_host.f90), user is passed back to them unchanged
     integer(c_int) function rp_pes_create_host(ndim, natom, potforce, hessian, user, out) &
          bind(C, name="rp_pes_create_host")
       import :: c_int, c_ptr, c_funptr
       integer(c_int), value :: ndim, natom
       type(c_funptr), value :: potforce, hessian
       type(c_ptr), value :: user
       type(c_ptr), intent(out) :: out
     end function rp_pes_create_host
     integer(c_int) function rp_pes_destroy(pes) bind(C, name="rp_pes_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: pes
     end function rp_pes_destroy
     integer(c_int) function rp_pes_set_v0(pes, v0) bind(C, name="rp_pes_set_v0")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: pes
       real(c_double), value :: v0
     end function rp_pes_set_v0
     integer(c_int) function rp_pes_potforce(pes, nbead, x, grad, energy) &
          bind(C, name="rp_pes_potforce")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: pes
       integer(c_int), value :: nbead
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(out) :: grad(*), energy(*)
     end function rp_pes_potforce
     integer(c_int) function rp_pes_hessian(pes, nbead, x, hess) bind(C, name="rp_pes_hessian")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: pes
       integer(c_int), value :: nbead
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(out) :: hess(*)
     end function rp_pes_hessian
     integer(c_int) function rp_pes_param_grad(pes, nbead, x, out, nparam) &
          bind(C, name="rp_pes_param_grad")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: pes
       integer(c_int), value :: nbead
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(out) :: out(*)
       integer(c_int), intent(out) :: nparam
     end function rp_pes_param_grad

     ! ---------------------------------------------------------------- plain path
     integer(c_int) function rp_nm_setup(pes, n, ndim, natom, mass, betan, tau, dt, gamma, &
          cayley, a, b, out) bind(C, name="rp_nm_setup")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: pes
       integer(c_int), value :: n, ndim, natom, cayley
       real(c_double), intent(in) :: mass(*), a(*), b(*)
       real(c_double), value :: betan, tau, dt, gamma
       type(c_ptr), intent(out) :: out
     end function rp_nm_setup
     integer(c_int) function rp_chain_destroy(c) bind(C, name="rp_chain_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: c
     end function rp_chain_destroy
     integer(c_int) function rp_chain_set_ends(c, a, b) bind(C, name="rp_chain_set_ends")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c
       real(c_double), intent(in) :: a(*), b(*)
     end function rp_chain_set_ends
     integer(c_int) function rp_chain_set_state(c, x, p) bind(C, name="rp_chain_set_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c
       real(c_double), intent(in) :: x(*), p(*)
     end function rp_chain_set_state
     integer(c_int) function rp_chain_get_state(c, x, p) bind(C, name="rp_chain_get_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c
       real(c_double), intent(out) :: x(*), p(*)
     end function rp_chain_get_state
     integer(c_int) function rp_chain_get_nm(c, T, lam, beadmass, beadvec) &
          bind(C, name="rp_chain_get_nm")
       import :: c_int, c_ptr
       type(c_ptr), value :: c, T, lam, beadmass, beadvec
     end function rp_chain_get_nm
     integer(c_int) function rp_chain_set_rng(c, seed, stream, cursor) &
          bind(C, name="rp_chain_set_rng")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: c
       integer(c_int64_t), value :: seed, stream, cursor
     end function rp_chain_set_rng
     integer(c_int) function rp_chain_get_cursor(c, cursor) bind(C, name="rp_chain_get_cursor")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: c
       integer(c_int64_t), intent(out) :: cursor
     end function rp_chain_get_cursor
     integer(c_int) function rp_chain_set_kick_cursor(c, kick_cursor) &
          bind(C, name="rp_chain_set_kick_cursor")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: c
       integer(c_int64_t), value :: kick_cursor
     end function rp_chain_set_kick_cursor
     integer(c_int) function rp_chain_get_kick_cursor(c, kick_cursor) &
          bind(C, name="rp_chain_get_kick_cursor")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: c
       integer(c_int64_t), intent(out) :: kick_cursor
     end function rp_chain_get_kick_cursor
     integer(c_int) function rp_nm_transform(c, direction, ncols, shift_dof, in, out) &
          bind(C, name="rp_nm_transform")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c
       integer(c_int), value :: direction, ncols
       integer(c_int), intent(in) :: shift_dof(*)
       real(c_double), intent(in) :: in(*)
       real(c_double), intent(out) :: out(*)
     end function rp_nm_transform
     integer(c_int) function rp_step_v(c, h) bind(C, name="rp_step_v")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c
       real(c_double), value :: h
     end function rp_step_v
     integer(c_int) function rp_step_nm(c, h) bind(C, name="rp_step_nm")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c
       real(c_double), value :: h
     end function rp_step_nm
     integer(c_int) function rp_step_langevin(c, noise) bind(C, name="rp_step_langevin")
       import :: c_int, c_ptr
       type(c_ptr), value :: c, noise
     end function rp_step_langevin
     integer(c_int) function rp_time_step_pile(c, noise) bind(C, name="rp_time_step_pile")
       import :: c_int, c_ptr
       type(c_ptr), value :: c, noise
     end function rp_time_step_pile
     integer(c_int) function rp_time_step_nm(c) bind(C, name="rp_time_step_nm")
       import :: c_int, c_ptr
       type(c_ptr), value :: c
     end function rp_time_step_nm
     integer(c_int) function rp_andersen_resample(c, noise) bind(C, name="rp_andersen_resample")
       import :: c_int, c_ptr
       type(c_ptr), value :: c, noise
     end function rp_andersen_resample
     integer(c_int) function rp_chain_set_path(c, npath, lampath, path) &
          bind(C, name="rp_chain_set_path")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c
       integer(c_int), value :: npath
       real(c_double), intent(in) :: lampath(*), path(*)
     end function rp_chain_set_path
     integer(c_int) function rp_init_path(c, xi, noise) bind(C, name="rp_init_path")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c, noise
       real(c_double), value :: xi
     end function rp_init_path
     integer(c_int) function rp_propagate_pile(c, nmc, imin, dhdrlimit, dbdl, xi, noise, &
          nnoise, dhdr, nreinit) bind(C, name="rp_propagate_pile")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: c, noise
       integer(c_int64_t), value :: nmc, imin, nnoise
       real(c_double), value :: dhdrlimit, xi
       real(c_double), intent(in) :: dbdl(*)
       real(c_double), intent(out) :: dhdr
       integer(c_int64_t), intent(out) :: nreinit
     end function rp_propagate_pile
     integer(c_int) function rp_propagate_nm(c, nmc, imin, noutput, dbdl, noise, nnoise, &
          kicks, nkicks, dhdr) bind(C, name="rp_propagate_nm")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: c, noise, kicks
       integer(c_int64_t), value :: nmc, imin, noutput, nnoise, nkicks
       real(c_double), intent(in) :: dbdl(*)
       real(c_double), intent(out) :: dhdr
     end function rp_propagate_nm

     ! ---------------------------------------------------------------- batched engine
     integer(c_int) function rp_batch_create(pes, mass, npoly, desc, tau, dt, gamma, &
          thermostat, noutput, dhdrlimit, cayley, seed, out) bind(C, name="rp_batch_create")
       import :: c_int, c_ptr, c_double, c_int64_t, rp_polymer_desc
       type(c_ptr), value :: pes
       real(c_double), intent(in) :: mass(*)
       integer(c_int), value :: npoly, thermostat, cayley
       type(rp_polymer_desc), intent(in) :: desc(*)
       real(c_double), value :: tau, dt, gamma, dhdrlimit
       integer(c_int64_t), value :: noutput, seed
       type(c_ptr), intent(out) :: out
     end function rp_batch_create
     integer(c_int) function rp_batch_destroy(b) bind(C, name="rp_batch_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: b
     end function rp_batch_destroy
     integer(c_int) function rp_batch_set_path(b, npath, lampath, path) &
          bind(C, name="rp_batch_set_path")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: b
       integer(c_int), value :: npath
       real(c_double), intent(in) :: lampath(*), path(*)
     end function rp_batch_set_path
     integer(c_int) function rp_batch_init_path(b, noise) bind(C, name="rp_batch_init_path")
       import :: c_int, c_ptr
       type(c_ptr), value :: b, noise
     end function rp_batch_init_path
     integer(c_int) function rp_batch_set_state(b, poly, x, p) bind(C, name="rp_batch_set_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: b
       integer(c_int), value :: poly
       real(c_double), intent(in) :: x(*), p(*)
     end function rp_batch_set_state
     integer(c_int) function rp_batch_get_state(b, poly, x, p) bind(C, name="rp_batch_get_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: b
       integer(c_int), value :: poly
       real(c_double), intent(out) :: x(*), p(*)
     end function rp_batch_get_state
     integer(c_int) function rp_batch_set_kicks(b, kicks, nkicks_per_poly) &
          bind(C, name="rp_batch_set_kicks")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: b, kicks
       integer(c_int64_t), value :: nkicks_per_poly
     end function rp_batch_set_kicks
     integer(c_int) function rp_batch_run(b, nsteps, imin, noise, nnoise_per_poly) &
          bind(C, name="rp_batch_run")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: b, noise
       integer(c_int64_t), value :: nsteps, imin, nnoise_per_poly
     end function rp_batch_run
     integer(c_int) function rp_batch_ti_readout(b, dhdr_sum, count, nreinit) &
          bind(C, name="rp_batch_ti_readout")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: b
       real(c_double), intent(out) :: dhdr_sum(*)
       integer(c_int64_t), intent(out) :: count(*), nreinit(*)
     end function rp_batch_ti_readout
     integer(c_int) function rp_batch_get_cursor(b, cursor) bind(C, name="rp_batch_get_cursor")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: b
       integer(c_int64_t), intent(out) :: cursor(*)
     end function rp_batch_get_cursor
     integer(c_int) function rp_batch_get_kick_cursor(b, kick_cursor) &
          bind(C, name="rp_batch_get_kick_cursor")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: b
       integer(c_int64_t), intent(out) :: kick_cursor(*)
     end function rp_batch_get_kick_cursor
     ! cursor, kick_cursor: c_loc of int64 arrays, or c_null_ptr to keep the current values
     integer(c_int) function rp_batch_set_cursor(b, cursor, kick_cursor) &
          bind(C, name="rp_batch_set_cursor")
       import :: c_int, c_ptr
       type(c_ptr), value :: b, cursor, kick_cursor
     end function rp_batch_set_cursor
     integer(c_int) function rp_batch_families(b, nfam, n, count) bind(C, name="rp_batch_families")
       import :: c_int, c_ptr
       type(c_ptr), value :: b, n, count
       integer(c_int), intent(out) :: nfam
     end function rp_batch_families
     integer(c_int) function rp_batch_set_launch(b, steps_per_launch, waves) &
          bind(C, name="rp_batch_set_launch")
       import :: c_int, c_ptr
       type(c_ptr), value :: b
       integer(c_int), value :: steps_per_launch, waves
     end function rp_batch_set_launch
     integer(c_int) function rp_batch_set_graphs(b, enable) bind(C, name="rp_batch_set_graphs")
       import :: c_int, c_ptr
       type(c_ptr), value :: b
       integer(c_int), value :: enable
     end function rp_batch_set_graphs
     integer(c_int) function rp_batch_autotune(b, trial_steps, best_steps, best_waves) &
          bind(C, name="rp_batch_autotune")
       import :: c_int, c_ptr
       type(c_ptr), value :: b
       integer(c_int), value :: trial_steps
       integer(c_int), intent(out) :: best_steps, best_waves
     end function rp_batch_autotune
     integer(c_int) function rp_batch_last_kernel_ms(b, avg_ms, launches) &
          bind(C, name="rp_batch_last_kernel_ms")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: b
       real(c_double), intent(out) :: avg_ms(*)
       integer(c_int64_t), intent(out) :: launches
     end function rp_batch_last_kernel_ms

     ! ---------------------------------------------------------------- instantons
     integer(c_int) function rp_ring_create(pes, mass, nchain, n, betan, a, b, fixedends, out) &
          bind(C, name="rp_ring_create")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: pes
       real(c_double), intent(in) :: mass(*), betan(*), a(*), b(*)
       integer(c_int), value :: nchain, fixedends
       integer(c_int), intent(in) :: n(*)
       type(c_ptr), intent(out) :: out
     end function rp_ring_create
     integer(c_int) function rp_ring_destroy(r) bind(C, name="rp_ring_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: r
     end function rp_ring_destroy
     integer(c_int) function rp_ring_set_x(r, x) bind(C, name="rp_ring_set_x")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       real(c_double), intent(in) :: x(*)
     end function rp_ring_set_x
     integer(c_int) function rp_ring_get_x(r, x) bind(C, name="rp_ring_get_x")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       real(c_double), intent(out) :: x(*)
     end function rp_ring_get_x
     integer(c_int) function rp_ring_um(r, energy, grad) bind(C, name="rp_ring_um")
       import :: c_int, c_ptr
       type(c_ptr), value :: r, energy, grad
     end function rp_ring_um
     integer(c_int) function rp_ring_hvp(r, v, hv) bind(C, name="rp_ring_hvp")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       real(c_double), intent(in) :: v(*)
       real(c_double), intent(out) :: hv(*)
     end function rp_ring_hvp
     integer(c_int) function rp_ring_hessian_band(r, chain, singlewell, band) &
          bind(C, name="rp_ring_hessian_band")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: chain, singlewell
       real(c_double), intent(out) :: band(*)
     end function rp_ring_hessian_band
     integer(c_int) function rp_ring_logdet(r, singlewell, skip_lowest, logdet_pos, nneg, eta1) &
          bind(C, name="rp_ring_logdet")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: singlewell, skip_lowest
       real(c_double), intent(out) :: logdet_pos(*), eta1(*)
       integer(c_int), intent(out) :: nneg(*)
     end function rp_ring_logdet
     integer(c_int) function rp_ring_lowest_eigs(r, singlewell, k, eigs) &
          bind(C, name="rp_ring_lowest_eigs")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       integer(c_int), value :: singlewell, k
       real(c_double), intent(out) :: eigs(*)
     end function rp_ring_lowest_eigs
     integer(c_int) function rp_ring_set_preconditioner(r, sigma) &
          bind(C, name="rp_ring_set_preconditioner")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       real(c_double), value :: sigma
     end function rp_ring_set_preconditioner
     integer(c_int) function rp_instanton(r, pgtol, factr, m, maxiter, x_inout, action, iters, &
          status) bind(C, name="rp_instanton")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       real(c_double), value :: pgtol, factr
       integer(c_int), value :: m, maxiter
       real(c_double), intent(inout) :: x_inout(*)
       real(c_double), intent(out) :: action(*)
       integer(c_int), intent(out) :: iters(*), status(*)
     end function rp_instanton
     integer(c_int) function rp_ring_param_grad(r, out, nparam) bind(C, name="rp_ring_param_grad")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: r
       real(c_double), intent(out) :: out(*)
       integer(c_int), intent(out) :: nparam
     end function rp_ring_param_grad
  end interface

contains

  ! A non-zero status prints the library's message and stops, as the reference does.
  subroutine rp_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    integer :: i, ierr
    character(len=512) :: msg
    if (rc == RP_OK) return
    ierr = rp_last_error(buf, 512)
    msg = ' '
    do i = 1, 512
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    end do
    write(*,*) trim(where), ': ', trim(msg)
    stop 1
  end subroutine rp_check

end module rp_hip
