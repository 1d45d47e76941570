! rp_selftest.f90 -- a Fortran caller of the device path through the drop-in modules
! (mcmod_mass, instantonmod, verletint of this directory), the way the reference's
! drivers call theirs.  Reads one case from stdin, prints the results after tagged lines.
!
! stdin:  kind ndim natom ndp nip / dp / ip (if nip > 0) / mass(natom)
!         n betan tau dt gamma nmc imin seed v0
!         a(ndof) / b(ndof) / dbdl(ndof) / x(n*ndof) / p(n*ndof)   (x, p as x(n,ndim,natom))
! stdout: PLUGIN  V, Vprime, Vdoubleprime at bead 1
!         PILE    x, p, dHdr after propagate_pimd_pile from (x, p), Philox seed, stream 0
!         PILE2   the same after a second init_nm, from the same (x, p), continuing the stream
!         UM      UMforceenergy energy and gradient at x with ends a, b
!         BAND    UMhessian(x, singlewell=.false.)
!         GAULEG  11 nodes and weights on [0, 1]
program rp_selftest
  use mcmod_mass
  use instantonmod
  use verletint
  implicit none
  integer:: kind, ndp, nip, nmc_in, imin_in, i
  integer(8):: seed
  double precision:: v0in, e, dHdr, xg(11), wg(11)
  double precision, allocatable:: dp(:), a(:,:), b(:,:), dbdl(:,:), x(:,:,:), p(:,:,:)
  double precision, allocatable:: x0(:,:,:), p0(:,:,:), g(:,:,:), band(:,:), grad1(:,:), hess1(:,:,:,:)
  integer, allocatable:: ip(:)

  read(*,*) kind, ndim, natom, ndp, nip
  allocate(dp(ndp), ip(nip), mass(natom))
  read(*,*) dp
  if (nip > 0) read(*,*) ip
  read(*,*) mass
  read(*,*) n, betan, tau, dt, gamma, nmc_in, imin_in, seed, v0in
  ndof = ndim*natom
  totdof = n*ndof
  allocate(a(ndim,natom), b(ndim,natom), dbdl(ndim,natom), x(n,ndim,natom), p(n,ndim,natom))
  read(*,*) a
  read(*,*) b
  read(*,*) dbdl
  read(*,*) x
  read(*,*) p
  call rp_plugin_setup(kind, ndim, natom, dp, ip)
  call V_init(0)
  V0 = v0in
  fixedends = .true.

  ! the plugin API on bead 1
  allocate(grad1(ndim,natom), hess1(ndim,natom,ndim,natom))
  write(*,'(A)') 'PLUGIN'
  write(*,'(ES25.17)') V(x(1,:,:))
  call Vprime(x(1,:,:), grad1)
  write(*,'(ES25.17)') grad1
  call Vdoubleprime(x(1,:,:), hess1)
  write(*,'(ES25.17)') hess1

  ! one PILE trajectory through verletint
  NMC = nmc_in
  imin = imin_in
  Noutput = 10
  rp_seed = seed
  call alloc_nm(0)
  call init_nm(a, b)
  allocate(x0(n,ndim,natom))
  x0 = x
  allocate(p0(n,ndim,natom))
  p0 = p
  call propagate_pimd_pile(x, p, a, b, dbdl, 0.0d0, dHdr)
  write(*,'(A)') 'PILE'
  write(*,'(ES25.17)') x
  write(*,'(ES25.17)') p
  write(*,'(ES25.17)') dHdr
  ! the next task of the process (pimd_par.f90:321-329): init_nm again, from the same
  ! start; its draws continue the stream, so it is not a replay of the first task
  call init_nm(a, b)
  x = x0
  p = p0
  call propagate_pimd_pile(x, p, a, b, dbdl, 0.0d0, dHdr)
  write(*,'(A)') 'PILE2'
  write(*,'(ES25.17)') x
  write(*,'(ES25.17)') p
  write(*,'(ES25.17)') dHdr
  call free_nm()

  ! the instanton routines at the initial chain
  allocate(g(n,ndim,natom), band(ndof+1, totdof))
  call UMforceenergy(x0, g, e, a, b)
  write(*,'(A)') 'UM'
  write(*,'(ES25.17)') e
  write(*,'(ES25.17)') g
  call UMhessian(x0, .false., band)
  write(*,'(A)') 'BAND'
  write(*,'(ES25.17)') band

  call gauleg(0.0d0, 1.0d0, xg, wg, 11)
  write(*,'(A)') 'GAULEG'
  do i = 1, 11
     write(*,'(2ES25.17)') xg(i), wg(i)
  end do
  write(*,'(A)') 'END'
end program rp_selftest
