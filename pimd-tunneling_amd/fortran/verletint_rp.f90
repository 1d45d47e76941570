! verletint_rp.f90 -- `module verletint` (verletmodule.f90) on the GPU.
!
! The reference drivers set this module's variables (NMC, Noutput, imin, dt, dHdrlimit,
! tau, gamma, cayley, restart, ...) and call alloc_nm, init_nm, init_path,
! propagate_pimd_pile / propagate_pimd_nm, write_restart and free_nm; the same names do the
! same here, each one a launch sequence on the device chain of libpimdtunnel.so.  The state
! arrays keep the reference layout x(n, ndim, natom) and are copied to and from the device
! per call.  The potential is reached through `module rp_bridge` (a device plugin, or any
! host plugin of the reference's API).
!
! Random numbers: the reference seeds MKL VSL MT19937 streams from the clock
! (verletmodule.f90:355-358) and draws from them for the whole run.  Here every process
! draws from one counter-based Philox stream (seed rp_seed, stream rp_stream = iproc):
! alloc_nm takes the seed from the clock as the reference does unless the driver has set
! rp_seed >= 0, and the stream positions (rp_cursor for the normals, rp_kick_cursor for the
! Andersen intervals) carry over from one task's init_nm to the next, so successive tasks of
! a process draw fresh numbers as the VSL streams do.
!
! Not carried over: the NR-FFT transform selected by use_mkl (the transform is the dense
! DST-I on the device either way) and the readhess initial momenta.
module verletint
  use iso_c_binding
  use rp_hip
  use mcmod_mass
  use rp_bridge
  use instantonmod
  implicit none

  integer::                         NMC, Noutput, imin, restart=0, iproc=0, restartnmc=0, restartunit
  double precision::                dt, dHdrlimit=-1.0d0
  double precision, allocatable::   transmatrix(:,:), beadvec(:,:), beadmass(:,:), lam(:)
  double precision, allocatable::   path(:,:,:), lampath(:), splinepath(:,:,:), Vpath(:)
  double precision, allocatable::   hesspath(:,:,:), splinehess(:,:,:)
  double precision::                tau=1.0d0, gamma=0.0d0
  logical::                         iprint=.false., use_mkl=.false., cayley=.false., readhess=.false.
  integer(c_int64_t)::              rp_seed=-1, rp_stream=0, rp_cursor=0, rp_kick_cursor=0
  type(c_ptr)::                     rp_chain_handle = c_null_ptr

contains

  subroutine alloc_nm(iproc_in)
    integer, intent(in):: iproc_in
    integer:: itime, irate, imax
    iproc = iproc_in
    if (rp_seed < 0) then
       call system_clock(itime, irate, imax)
       rp_seed = int(mod(itime + 5*iproc, 1000), c_int64_t)
    end if
    rp_stream = int(iproc, c_int64_t)
    rp_cursor = 0
    rp_kick_cursor = 0
    write(*,*) "proc", iproc, "running with Philox seed", rp_seed, "stream", rp_stream
  end subroutine alloc_nm

  ! the device chain's stream position is kept when the chain goes away
  subroutine drop_chain(where)
    character(len=*), intent(in):: where
    if (.not. c_associated(rp_chain_handle)) return
    call rp_check(rp_chain_get_cursor(rp_chain_handle, rp_cursor), where)
    call rp_check(rp_chain_get_kick_cursor(rp_chain_handle, rp_kick_cursor), where)
    call rp_check(rp_chain_destroy(rp_chain_handle), where)
    rp_chain_handle = c_null_ptr
  end subroutine drop_chain

  ! init_nm (verletmodule.f90:306-338) plus the Langevin table of :380-386
  subroutine init_nm(a, b)
    double precision, intent(in):: a(:,:), b(:,:)
    double precision, allocatable, target:: T(:,:), lm(:), bm(:,:), bv(:,:)
    call drop_chain('init_nm')
    call rp_check(rp_nm_setup(rp_bridge_pes(), int(n, c_int), int(ndim, c_int), int(natom, c_int), &
         mass, betan, tau, dt, gamma, merge(1_c_int, 0_c_int, cayley), a, b, rp_chain_handle), &
         'init_nm')
    call rp_check(rp_chain_set_rng(rp_chain_handle, rp_seed, rp_stream, rp_cursor), 'init_nm')
    call rp_check(rp_chain_set_kick_cursor(rp_chain_handle, rp_kick_cursor), 'init_nm')
    allocate(T(n, n), lm(n), bm(natom, n), bv(n, ndof))
    call rp_check(rp_chain_get_nm(rp_chain_handle, c_loc(T), c_loc(lm), c_loc(bm), c_loc(bv)), &
         'init_nm')
    if (allocated(transmatrix)) deallocate(transmatrix, beadvec, beadmass, lam)
    call move_alloc(T, transmatrix)   ! symmetric, so the C row-major copy is the matrix
    call move_alloc(lm, lam)
    call move_alloc(bm, beadmass)     ! beadmass(atom, mode)
    call move_alloc(bv, beadvec)      ! beadvec(mode, dof)
    if (allocated(lampath)) call rp_check(rp_chain_set_path(rp_chain_handle, &
         int(size(lampath), c_int), lampath, path), 'init_nm')
  end subroutine init_nm

  subroutine free_nm()
    call drop_chain('free_nm')
    if (allocated(transmatrix)) deallocate(transmatrix, beadvec, beadmass, lam)
  end subroutine free_nm

  subroutine put_state(x, p)
    double precision, intent(in):: x(:,:,:), p(:,:,:)
    call rp_check(rp_chain_set_state(rp_chain_handle, x, p), 'verletint')
  end subroutine put_state

  subroutine get_state(x, p)
    double precision, intent(out):: x(:,:,:), p(:,:,:)
    call rp_check(rp_chain_get_state(rp_chain_handle, x, p), 'verletint')
  end subroutine get_state

  ! init_path (verletmodule.f90:32-119): beads along the path spline, NM momenta
  subroutine init_path(xi, x, p)
    double precision:: xi, x(:,:,:), p(:,:,:)
    if (.not. allocated(lampath)) then
       write(*,*) "init_path: no path (set lampath and path before init_nm)"
       stop 1
    end if
    call rp_check(rp_init_path(rp_chain_handle, xi, c_null_ptr), 'init_path')
    call get_state(x, p)
  end subroutine init_path

  ! write_restart (verletmodule.f90:162-185): positions with the running TI sum, then
  ! momenta with the step count, one block of natom lines per bead, on restartunit
  subroutine write_restart(xprop, pprop, ii, dHdr)
    double precision, intent(in):: xprop(:,:,:), pprop(:,:,:), dHdr
    integer, intent(in):: ii
    integer:: i, j
    rewind(restartunit)
    do i = 1, n
       write(restartunit,*) natom
       write(restartunit,*) dHdr
       do j = 1, natom
          write(restartunit,*) label(j), xprop(i,1,j), xprop(i,2,j), xprop(i,3,j)
       end do
    end do
    do i = 1, n
       write(restartunit,*) natom
       write(restartunit,*) ii
       do j = 1, natom
          write(restartunit,*) label(j), pprop(i,1,j), pprop(i,2,j), pprop(i,3,j)
       end do
    end do
  end subroutine write_restart

  ! propagate_pimd_pile (verletmodule.f90:372-416).  dHdr is the running TI sum on entry
  ! when resuming (restart = 2, pimd_par.f90:355-370) and the mean on return.  With
  ! restart > 0 the trajectory runs in device segments between the reference's output
  ! points (every Noutput steps after imin, :389-395), where write_restart records the
  ! state before that step; the segments continue one stream, so the trajectory is the
  ! same as one uninterrupted call.
  subroutine propagate_pimd_pile(xprop, vprop, a, b, dbdl, xi, dHdr)
    double precision:: xprop(:,:,:), vprop(:,:,:), a(:,:), b(:,:), dbdl(:,:), xi, dHdr
    integer(c_int64_t):: nre, nretot
    integer:: ii, done, cnt
    double precision:: seg, total
    total = 0.0d0
    if (restart >= 2) total = dHdr
    nretot = 0
    done = 0
    cnt = 0
    call put_state(xprop, vprop)
    if (restart > 0 .and. Noutput > 0) then
       do ii = 1, NMC
          cnt = cnt + 1
          if (cnt >= Noutput .and. ii > imin) then
             cnt = 0
             call run_segment(ii - 1 - done)
             call get_state(xprop, vprop)
             call write_restart(xprop, vprop, ii + restartnmc, total)
          end if
       end do
    end if
    call run_segment(NMC - done)
    if (nretot > 0) write(*,*) "Over limit: path re-initialised", nretot, "times"
    call get_state(xprop, vprop)
    dHdr = total/dble(NMC + restartnmc - imin)
  contains
    ! steps done+1 .. done+nseg of the trajectory; TI after global step imin
    subroutine run_segment(nseg)
      integer, intent(in):: nseg
      integer:: imloc
      if (nseg <= 0) return
      imloc = min(max(imin - done, 0), nseg)
      call rp_check(rp_propagate_pile(rp_chain_handle, int(nseg, c_int64_t), int(imloc, c_int64_t), &
           dHdrlimit, dbdl, xi, c_null_ptr, 0_c_int64_t, seg, nre), 'propagate_pimd_pile')
      if (nseg > imloc) total = total + seg*dble(nseg - imloc)
      nretot = nretot + nre
      done = done + nseg
    end subroutine run_segment
  end subroutine propagate_pimd_pile

  ! propagate_pimd_nm (verletmodule.f90:190-250), the Andersen thermostat
  subroutine propagate_pimd_nm(xprop, vprop, a, b, dbdl, dHdr)
    double precision:: xprop(:,:,:), vprop(:,:,:), a(:,:), b(:,:), dbdl(:,:), dHdr
    call put_state(xprop, vprop)
    call rp_check(rp_propagate_nm(rp_chain_handle, int(NMC, c_int64_t), int(imin, c_int64_t), &
         int(Noutput, c_int64_t), dbdl, c_null_ptr, 0_c_int64_t, c_null_ptr, 0_c_int64_t, dHdr), &
         'propagate_pimd_nm')
    call get_state(xprop, vprop)
  end subroutine propagate_pimd_nm

  subroutine time_step_pile(xprop, pprop)        ! :423-435
    double precision, intent(inout):: xprop(:,:,:), pprop(:,:,:)
    call put_state(xprop, pprop)
    call rp_check(rp_time_step_pile(rp_chain_handle, c_null_ptr), 'time_step_pile')
    call get_state(xprop, pprop)
  end subroutine time_step_pile

  subroutine time_step_nm(xprop, pprop)          ! :291-302
    double precision, intent(inout):: xprop(:,:,:), pprop(:,:,:)
    call put_state(xprop, pprop)
    call rp_check(rp_time_step_nm(rp_chain_handle), 'time_step_nm')
    call get_state(xprop, pprop)
  end subroutine time_step_nm

  subroutine step_nm(time, x, p, transform)      ! :494-561
    double precision:: time, x(:,:,:), p(:,:,:)
    logical:: transform
    call put_state(x, p)
    call rp_check(rp_step_nm(rp_chain_handle, time), 'step_nm')
    call get_state(x, p)
  end subroutine step_nm

  subroutine step_v(time, x, p, force, recalculate)  ! :565-585
    double precision:: time, x(:,:,:), p(:,:,:), force(:,:,:)
    logical:: recalculate
    double precision:: p0(size(p,1), size(p,2), size(p,3))
    p0 = p
    call put_state(x, p)
    call rp_check(rp_step_v(rp_chain_handle, time), 'step_v')
    call get_state(x, p)
    force = (p0 - p)/time   ! the gradient the kick used, as the reference leaves it
  end subroutine step_v

  subroutine step_langevin(pprop)                ! :634-667
    double precision, intent(inout):: pprop(:,:,:)
    double precision, dimension(size(pprop,1), size(pprop,2), size(pprop,3)):: x, pdev
    call get_state(x, pdev)   ! the device positions stay as they are
    call put_state(x, pprop)
    call rp_check(rp_step_langevin(rp_chain_handle, c_null_ptr), 'step_langevin')
    call get_state(x, pprop)
  end subroutine step_langevin

  ! nmtransform_forward / _backward (:254-286) of one n-vector; bead > 0 names the dof
  ! whose beadvec shift applies (positions), 0 none (momenta)
  subroutine nmtransform_forward(xprop, qprop, bead)
    double precision, intent(in):: xprop(:)
    double precision, intent(out):: qprop(:)
    integer, intent(in):: bead
    integer(c_int):: sh(1)
    sh(1) = int(bead - 1, c_int)
    if (bead <= 0) sh(1) = -1
    call rp_check(rp_nm_transform(rp_chain_handle, 0_c_int, 1_c_int, sh, xprop, qprop), &
         'nmtransform_forward')
  end subroutine nmtransform_forward

  subroutine nmtransform_backward(qprop, xprop, bead)
    double precision, intent(in):: qprop(:)
    double precision, intent(out):: xprop(:)
    integer, intent(in):: bead
    integer(c_int):: sh(1)
    sh(1) = int(bead - 1, c_int)
    if (bead <= 0) sh(1) = -1
    call rp_check(rp_nm_transform(rp_chain_handle, 1_c_int, 1_c_int, sh, qprop, xprop), &
         'nmtransform_backward')
  end subroutine nmtransform_backward

  ! Gauss-Legendre nodes and weights on [x1, x2] (verletmodule.f90:124-160): Newton on
  ! the Legendre recurrence from the Chebyshev-like first guess
  subroutine gauleg(x1, x2, x, w, nintegral)
    integer, intent(in):: nintegral
    double precision, intent(in):: x1, x2
    double precision, intent(out):: x(nintegral), w(nintegral)
    double precision, parameter:: eps = 3.0d-14
    double precision:: z, z1, p1, p2, p3, pp, xm, xl
    integer:: i, j, m
    m = (nintegral + 1)/2
    xm = 0.5d0*(x2 + x1)
    xl = 0.5d0*(x2 - x1)
    do i = 1, m
       z = cos(pi*(i - 0.25d0)/(nintegral + 0.5d0))
       do
          p1 = 1.0d0
          p2 = 0.0d0
          do j = 1, nintegral
             p3 = p2
             p2 = p1
             p1 = ((2.0d0*j - 1.0d0)*z*p2 - (j - 1.0d0)*p3)/j
          end do
          pp = nintegral*(z*p1 - p2)/(z*z - 1.0d0)
          z1 = z
          z = z1 - p1/pp
          if (abs(z - z1) <= eps) exit
       end do
       x(i) = xm - xl*z
       x(nintegral + 1 - i) = xm + xl*z
       w(i) = 2.0d0*xl/((1.0d0 - z*z)*pp*pp)
       w(nintegral + 1 - i) = w(i)
    end do
  end subroutine gauleg

end module verletint
