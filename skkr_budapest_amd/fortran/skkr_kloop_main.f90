! Fortran host program: reads a system (structure constants, t^-1, k-mesh,
! screening), runs the k-loop on the GPU through the C-ABI and writes the
! BZ-summed tau and the per-k l-resolved traces of the physical tau.
!   skkr_kloop <input> <output> [nprocs]
! Input (stream, little endian): int32 lk,s,P,nR,nk,itermax,nl; real*8 errmax,
! e_re, e_im; real*8 R(2,nR); complex*16 gintra(n,n,nR), ginter(n,n,nR),
! tinv(lk,lk,P*s), tinvl(lk,lk,s), tinvr(lk,lk,s); real*8 k(2,nk), w(nk);
! int32 l_of_index(lk); complex*16 alpha(lk,P*s).
! With nprocs > 1 the k-points are split round-robin over emulated ranks
! (modulo(ik-1,nprocs)==myrank) and the partial sums added, as the reference's
! MPI_Allreduce does.  Output: complex*16 tau(lk,lk,P*s),
! taukdiag(nk,0:nl-1,P*s).
program skkr_kloop_main
  use iso_c_binding
  use skkr_gpu
  use kloop_gpu
  implicit none
  integer(c_int) :: lk, s, np, nr, nk, itmax, nl, rc
  real(c_double) :: errmax, ere, eim
  real(c_double), allocatable :: R(:), k(:,:), w(:), xk(:,:)
  integer(c_int), allocatable :: lidx(:)
  complex(c_double_complex), allocatable :: gi(:), go(:), ti(:), tl(:), tr(:), al(:)
  complex(8), allocatable :: tau(:,:,:), tkd(:,:,:)
  type(skkr_config) :: cfg
  type(c_ptr) :: ctx
  character(len=512) :: fin, fout, arg
  integer :: n, nprocs, myrank
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  nprocs = 1
  if (command_argument_count() >= 3) then
     call get_command_argument(3, arg)
     read(arg, *) nprocs
  end if
  open(11, file=fin, access='stream', form='unformatted', status='old')
  read(11) lk, s, np, nr, nk, itmax, nl
  read(11) errmax, ere, eim
  n = lk * s
  allocate(R(2 * nr), gi(n * n * nr), go(n * n * nr), ti(lk * lk * np * s))
  allocate(tl(lk * lk * s), tr(lk * lk * s), k(2, nk), w(nk), xk(nk, 2))
  allocate(lidx(lk), al(lk * np * s))
  read(11) R, gi, go, ti, tl, tr, k, w, lidx, al
  close(11)
  xk(:, 1) = k(1, :)
  xk(:, 2) = k(2, :)
  cfg = skkr_config(lk, s, np, errmax, itmax, itmax, SKKR_FP64, 0.02d0, 0, 0)
  rc = skkr_create(ctx, cfg)
  if (rc /= SKKR_OK) stop 'skkr_create failed'
  rc = skkr_set_structure(ctx, nr, R, gi, go)
  if (rc /= SKKR_OK) stop 'skkr_set_structure failed'
  rc = skkr_set_tinv(ctx, ti, tl, tr)
  if (rc /= SKKR_OK) stop 'skkr_set_tinv failed'
  rc = skkr_set_screening(ctx, nl, lidx, al)
  if (rc /= SKKR_OK) stop 'skkr_set_screening failed'
  allocate(tau(lk, lk, np * s), tkd(nk, 0:nl - 1, np * s))
  tau = (0.d0, 0.d0)
  tkd = (0.d0, 0.d0)
  do myrank = 0, nprocs - 1
     call cpacoord_kloop_gpu(ctx, cmplx(ere, eim, 8), nk, xk, w, nk, lk, lk, np * s, np * s, &
          nl, nl - 1, tau, tkd, nprocs, myrank)
  end do
  call skkr_destroy(ctx)
  open(12, file=fout, access='stream', form='unformatted', status='replace')
  write(12) tau
  write(12) tkd
  close(12)
end program skkr_kloop_main
