! Host program in the reference's language calling the structure-constant
! producer through the ISO_C_BINDING module: the per-k gstore calls of
! cpacoord (nrsp-scf/cpacoord.f:205) replaced by one skkr_gstore_batch.
!
! Input (stream): int32 lmax, s, nextra, nk; real*8 eta, a1(2), a2(2), E(re,im);
! real*8 cvec(3, s+1); complex*16 alpha(0:lmax, s); real*8 park(2, nk).
! Output (stream): complex*16 g2d(nl2, nl2, -s:s, s, 0:2) per k.
program skkr_gstore_main
  use iso_c_binding
  use skkr_gpu
  implicit none
  integer(c_int) :: lmax, s, nextra, nk, nl2, ierr
  real(c_double) :: eta, a1(2), a2(2), e(2)
  real(c_double), allocatable :: cvec(:,:), park(:,:)
  complex(c_double_complex), allocatable :: alpha(:,:), g2d(:,:,:,:,:,:)
  integer(c_int), allocatable, target :: st(:)
  type(c_ptr) :: sc
  type(skkr_str_config) :: cfg
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(11, file=fin, access='stream', form='unformatted', status='old')
  read(11) lmax, s, nextra, nk
  read(11) eta, a1, a2, e
  nl2 = (lmax + 1)**2
  allocate(cvec(3, s + 1), alpha(0:lmax, s), park(2, nk))
  allocate(g2d(nl2, nl2, -s:s, s, 0:2, nk), st(nk))
  read(11) cvec
  read(11) alpha
  read(11) park
  close(11)

  cfg = skkr_str_config(lmax, s, nextra, eta, a1, a2, 0, 0)
  ierr = skkr_str_create(sc, cfg, cvec)
  if (ierr /= SKKR_OK) then
    write(*, *) 'skkr_str_create failed', ierr
    stop 1
  end if
  ierr = skkr_gstore_batch(sc, e(1), e(2), alpha, nk, park, g2d, c_loc(st))
  if (ierr /= SKKR_OK) then
    write(*, *) 'skkr_gstore_batch failed', ierr
    stop 1
  end if
  call skkr_str_destroy(sc)

  open(12, file=fout, access='stream', form='unformatted', status='replace')
  write(12) g2d
  close(12)
end program skkr_gstore_main
