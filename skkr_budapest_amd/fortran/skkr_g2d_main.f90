! k-loop of a bulk-like stack written against the reference's own
! subroutine interfaces: per k-point it scatters G(k) into the layer-indexed
! g2d(lm, lm', j, i, plane) array that packer.f reads (planes 0..2: left
! lead, interface, right lead), and calls dugo_gpu('L'), dugo_gpu('R') and
! tau2d_gpu with the argument lists of main/dugo.f:9-10 and main/tau2d.f:2-3,
! all arrays padded past what the system needs.
!
! Input (stream): int32 lk, s, P, nk, itermax, nsave; real*8 errmax;
! int32 isave(nsave); complex*16 tinv(lk,lk,P*s), tinvl(lk,lk,s),
! tinvr(lk,lk,s); real*8 w(nk); per k complex*16 gintra(n,n), ginter(n,n)
! (n = lk*s, G itself).
! Output (stream): complex*16 taubz(lk,lk,P*s), trk(P*s,nk),
! tausave(lk,lk,P*s,nsave); int32 iters(2,nk).
program skkr_g2d_main
  use iso_c_binding
  use skkr_gpu
  use skkr_dropin
  implicit none
  integer(c_int) :: lk, s, np, nk, itmx, nsave, ierr
  real(c_double) :: errm
  integer :: n, ns, ik, li, l, isv, lmax
  integer :: lmmaxp, minprc, mprc1, mintfc, lkmaxp, mdimp, ipl
  integer(c_int), allocatable :: isave(:), iters(:,:), ninprc(:)
  complex(c_double_complex), allocatable :: tinv(:,:,:), tinvl(:,:,:), tinvr(:,:,:)
  complex(c_double_complex), allocatable :: gi(:,:), go(:,:)
  complex(c_double_complex), allocatable :: g2d(:,:,:,:,:), tminv(:,:,:), tminvl(:,:,:), tminvr(:,:,:)
  complex(c_double_complex), allocatable :: xx(:,:), yy(:,:), tau(:,:,:)
  complex(c_double_complex), allocatable :: taubz(:,:,:), trk(:,:), tsave(:,:,:,:)
  real(c_double), allocatable :: w(:)
  logical :: conv
  type(c_ptr) :: ctx
  type(skkr_config) :: cfg
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(11, file=fin, access='stream', form='unformatted', status='old')
  read(11) lk, s, np, nk, itmx, nsave
  read(11) errm
  allocate(isave(nsave))
  read(11) isave
  n = lk * s
  ns = np * s
  ! the reference's param.h role: every dimension past what the system needs
  lmmaxp = lk + 2
  minprc = s + 1
  mprc1 = np + 1
  mintfc = ns + 3
  lkmaxp = lk + 1
  mdimp = n + 5
  lmax = 0  ! not read for irel = 0
  allocate(tinv(lk, lk, ns), tinvl(lk, lk, s), tinvr(lk, lk, s), w(nk))
  read(11) tinv, tinvl, tinvr
  read(11) w
  allocate(gi(n, n), go(n, n))
  allocate(g2d(lmmaxp, lmmaxp, -minprc:minprc, minprc, 0:mprc1 + 1))
  allocate(tminv(lkmaxp, lkmaxp, mintfc), tminvl(lkmaxp, lkmaxp, minprc), tminvr(lkmaxp, lkmaxp, minprc))
  allocate(xx(mdimp, mdimp), yy(mdimp, mdimp), tau(lkmaxp, lkmaxp, mintfc))
  allocate(taubz(lk, lk, ns), trk(ns, nk), tsave(lk, lk, ns, nsave), iters(2, nk), ninprc(0:np + 1))
  tminv = (0.0d0, 0.0d0)
  tminvl = (0.0d0, 0.0d0)
  tminvr = (0.0d0, 0.0d0)
  tminv(1:lk, 1:lk, 1:ns) = tinv
  tminvl(1:lk, 1:lk, 1:s) = tinvl
  tminvr(1:lk, 1:lk, 1:s) = tinvr
  ninprc = s

  cfg = skkr_config(lk, s, np, errm, itmx, itmx, SKKR_FP64, 0.02d0, 0, 0)
  ierr = skkr_create(ctx, cfg)
  if (ierr /= SKKR_OK) then
    write(*, *) 'skkr_create failed', ierr
    stop 1
  end if
  call skkr_dropin_init(ctx, lmmaxp, minprc, mprc1, mintfc, np, ninprc, (1.0d0, 0.0d0))

  taubz = (0.0d0, 0.0d0)
  do ik = 1, nk
    read(11) gi
    read(11) go
    g2d = (0.0d0, 0.0d0)
    do ipl = 0, 2
      call fill_plane(g2d(:, :, :, :, ipl))
    end do
    call dugo_gpu(lmax, lk, tminvl, lkmaxp, g2d(1, 1, -minprc, 1, 0), g2d(1, 1, -minprc, 1, 1), yy, &
         mdimp, 'L', .true., 0, conv)
    if (.not. conv) stop 3
    iters(1, ik) = dropin_last_iters
    call dugo_gpu(lmax, lk, tminvr, lkmaxp, g2d(1, 1, -minprc, 1, 2), g2d(1, 1, -minprc, 1, 1), xx, &
         mdimp, 'R', .true., 0, conv)
    if (.not. conv) stop 3
    iters(2, ik) = dropin_last_iters
    call tau2d_gpu(0, 1.0d0, .true., lmax, lk, lkmaxp, ns, tminv, g2d, xx, yy, tau, mdimp)
    do li = 1, ns
      trk(li, ik) = (0.0d0, 0.0d0)
      do l = 1, lk
        trk(li, ik) = trk(li, ik) + tau(l, l, li)
      end do
      taubz(:, :, li) = taubz(:, :, li) + w(ik) * tau(1:lk, 1:lk, li)
    end do
    do isv = 1, nsave
      if (isave(isv) == ik) tsave(:, :, :, isv) = tau(1:lk, 1:lk, 1:ns)
    end do
  end do
  close(11)
  call skkr_destroy(ctx)

  open(12, file=fout, access='stream', form='unformatted', status='replace')
  write(12) taubz, trk, tsave
  write(12) iters
  close(12)

contains

  ! one principal layer's window from G(k): the layer pairs inside the PL
  ! from gintra, the couplings to the PL before from ginter^dagger and to
  ! the PL after from ginter, blocks (j, i) as packer.f indexes them
  subroutine fill_plane(g2p)
    complex(c_double_complex), intent(inout) :: g2p(lmmaxp, lmmaxp, -minprc:minprc, minprc)
    integer :: i, j, jp
    do i = 1, s
      do j = -s, -i
        jp = s + i + j
        g2p(1:lk, 1:lk, j, i) = conjg(transpose(go((jp - 1) * lk + 1:jp * lk, (i - 1) * lk + 1:i * lk)))
      end do
      do j = 1 - i, s - i
        jp = i + j
        g2p(1:lk, 1:lk, j, i) = gi((i - 1) * lk + 1:i * lk, (jp - 1) * lk + 1:jp * lk)
      end do
      do j = s - i + 1, s
        jp = j + i - s
        g2p(1:lk, 1:lk, j, i) = go((i - 1) * lk + 1:i * lk, (jp - 1) * lk + 1:jp * lk)
      end do
    end do
  end subroutine fill_plane

end program skkr_g2d_main
