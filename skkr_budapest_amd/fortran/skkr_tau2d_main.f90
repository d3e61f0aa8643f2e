! Host program in the reference's language replacing the tau2d calls of a
! k-loop (main/tau2d.f:2) by one batched call through the ISO_C_BINDING
! module: skkr_tau2d_batch for a bulk-like stack (bulkgeo) or
! skkr_tau2d_layered_batch for a layered one (bulkgeo false, one set of
! packer blocks per principal layer, tau2d.f:55-56,90-91).
!
! Input (stream): int32 lk, s, P, nb, layered; complex*16 tinv(lk,lk,P*s),
! tinvl(lk,lk,s), tinvr(lk,lk,s); m00, m01, m10 as (n,n,P,nb) if layered
! else (n,n,nb), n = lk*s (the packer blocks); x(n,n,nb), y(n,n,nb) (the
! dugo results); real*8 wgeff.
! Output (stream): complex*16 tau(lk,lk,P*s,nb); int32 status(nb).
program skkr_tau2d_main
  use iso_c_binding
  use skkr_gpu
  implicit none
  integer(c_int) :: lk, s, np, nb, layered, n, npl, ierr
  real(c_double) :: wgeff
  complex(c_double_complex), allocatable :: tinv(:,:,:), tinvl(:,:,:), tinvr(:,:,:)
  complex(c_double_complex), allocatable :: m00(:,:,:,:), m01(:,:,:,:), m10(:,:,:,:)
  complex(c_double_complex), allocatable :: x(:,:,:), y(:,:,:), tau(:,:,:,:)
  integer(c_int), allocatable :: st(:)
  type(c_ptr) :: ctx
  type(skkr_config) :: cfg
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(11, file=fin, access='stream', form='unformatted', status='old')
  read(11) lk, s, np, nb, layered
  n = lk * s
  npl = 1
  if (layered /= 0) npl = np
  allocate(tinv(lk, lk, np * s), tinvl(lk, lk, s), tinvr(lk, lk, s))
  allocate(m00(n, n, npl, nb), m01(n, n, npl, nb), m10(n, n, npl, nb))
  allocate(x(n, n, nb), y(n, n, nb), tau(lk, lk, np * s, nb), st(nb))
  read(11) tinv, tinvl, tinvr
  read(11) m00, m01, m10
  read(11) x, y
  read(11) wgeff
  close(11)

  cfg = skkr_config(lk, s, np, 1.0d-8, 50, 50, SKKR_FP64, 0.02d0, 0, 0)
  ierr = skkr_create(ctx, cfg)
  if (ierr /= SKKR_OK) then
    write(*, *) 'skkr_create failed', ierr
    stop 1
  end if
  ierr = skkr_set_tinv(ctx, tinv, tinvl, tinvr)
  if (ierr /= SKKR_OK) then
    write(*, *) 'skkr_set_tinv failed', ierr
    stop 1
  end if
  if (layered /= 0) then
    ierr = skkr_tau2d_layered_batch(ctx, nb, m00, m01, m10, x, y, wgeff, tau, st)
  else
    ierr = skkr_tau2d_batch(ctx, nb, m00, m01, m10, x, y, wgeff, tau, st)
  end if
  call skkr_destroy(ctx)
  if (ierr /= SKKR_OK) then
    write(*, *) 'tau2d failed', ierr
    stop 1
  end if

  open(12, file=fout, access='stream', form='unformatted', status='replace')
  write(12) tau
  write(12) st
  close(12)
end program skkr_tau2d_main
