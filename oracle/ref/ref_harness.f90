! ref_harness.f90 -- drives the reference's instantonmod routines (UM, UMprime,
! UMforceenergy, UMhessian, instanton) on inputs read from stdin and prints the
! results after a RESULT line.  TEST INFRASTRUCTURE ONLY (built into oracle/_ref/).
!
! stdin:  task (um | hessian | instanton)
!         kind ndim natom n ndp nip
!         dp(ndp)  ip(nip)  mass(natom)
!         betan fixedends(0/1) singlewell(0/1) v0 eps2
!         a(ndof)  b(ndof)  x(n*ndof, column-major (n, ndim, natom))
! stdout: RESULT, then for um: UM(x), then UMforceenergy's energy and gradient, then
!         UMprime; for hessian: the band (ndof+1, n*ndof); for instanton: U_M at the
!         optimum and the optimised x.
program ref_harness
  use mcmod_mass
  use instantonmod
  implicit none
  character(len=16):: task
  integer:: kind_, ndp, nip, ifixed, isingle, i
  double precision:: v0in, e, e2
  double precision, allocatable:: x(:,:,:), g(:,:,:), g2(:,:,:), a(:,:), b(:,:), band(:,:)

  read(*,*) task
  read(*,*) kind_, ndim, natom, n, ndp, nip
  pes_kind = kind_
  allocate(pes_dp(ndp), pes_ip(max(nip,1)), mass(natom))
  read(*,*) pes_dp
  if (nip > 0) read(*,*) pes_ip(1:nip)
  read(*,*) mass
  read(*,*) betan, ifixed, isingle, v0in, eps2
  ndof = ndim*natom
  totdof = n*ndof
  call V_init(0)
  V0 = v0in
  fixedends = ifixed /= 0
  allocate(a(ndim,natom), b(ndim,natom), x(n,ndim,natom), g(n,ndim,natom), g2(n,ndim,natom))
  allocate(label(natom))
  label = 'X'
  read(*,*) a
  read(*,*) b
  read(*,*) x
  write(*,'(A)') 'RESULT'
  select case (trim(task))
  case ('um')
     e = UM(x, a, b)
     call UMforceenergy(x, g, e2, a, b)
     call UMprime(x, g2, a, b)
     write(*,'(ES25.17)') e
     write(*,'(ES25.17)') e2
     write(*,'(ES25.17)') g
     write(*,'(ES25.17)') g2
  case ('hessian')
     allocate(band(ndof+1, totdof))
     call UMhessian(x, isingle /= 0, band)
     write(*,'(ES25.17)') band
  case ('instanton')
     call instanton(x, a, b)
     write(*,'(A)') 'OPTIMUM'
     write(*,'(ES25.17)') UM(x, a, b)
     write(*,'(ES25.17)') x
  case default
     write(*,*) 'unknown task ', task
     stop
  end select
end program ref_harness
