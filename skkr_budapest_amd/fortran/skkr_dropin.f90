! tau2d and dugo with the reference's own argument lists, on the GPU:
!   call tau2d_gpu(irel,wgeff,bulkgeo,lmax,lkmax,lkmaxp,nintfc,tminv,g2d,x,y,tau,mdimp)
!   call dugo_gpu(lmax,lkmax,tminvb,lkmaxp,g2d,g2di,xy,mdimp,side,bulk,irel,sgfconv)
! take exactly the arguments of skkr_kpar_v6/main/tau2d.f:2-3 and
! main/dugo.f:9-10.  What those routines read from param.h (lmmaxp, minprc,
! mprc1, mintfc) and from the common blocks /lay2d/ (nprc, ninprc), /relfac/
! (fac) is handed over once by skkr_dropin_init, together with the context
! (lk = lkmax, sites_per_pl = ninprc(1), n_pl = nprc; its errmax and itermax
! play /iterpar/).  Errors stop the program, as the reference's routines do.
module skkr_dropin
  use iso_c_binding
  use skkr_gpu
  implicit none
  private
  public :: skkr_dropin_init, tau2d_gpu, dugo_gpu, dropin_last_iters

  type(c_ptr), save :: dctx = c_null_ptr
  integer(c_int), save :: d_lmmaxp = 0, d_minprc = 0, d_mprc1 = 0, d_mintfc = 0, d_nprc = 0
  integer(c_int), allocatable, save :: d_ninprc(:)
  complex(c_double_complex), save :: d_fac = (1.0d0, 0.0d0)
  integer(c_int), save :: dropin_last_iters = 0

contains

  subroutine skkr_dropin_init(ctx, lmmaxp, minprc, mprc1, mintfc, nprc, ninprc, fac)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: lmmaxp, minprc, mprc1, mintfc, nprc
    integer, intent(in) :: ninprc(0:nprc + 1)
    complex(c_double_complex), intent(in) :: fac
    dctx = ctx
    d_lmmaxp = lmmaxp
    d_minprc = minprc
    d_mprc1 = mprc1
    d_mintfc = mintfc
    d_nprc = nprc
    if (allocated(d_ninprc)) deallocate(d_ninprc)
    allocate(d_ninprc(0:nprc + 1))
    d_ninprc = ninprc
    d_fac = fac
  end subroutine skkr_dropin_init

  subroutine tau2d_gpu(irel, wgeff, bulkgeo, lmax, lkmax, lkmaxp, nintfc, tminv, g2d, x, y, tau, mdimp)
    integer, intent(in) :: irel, lmax, lkmax, lkmaxp, nintfc, mdimp
    real(c_double), intent(in) :: wgeff
    logical, intent(in) :: bulkgeo
    complex(c_double_complex), intent(in) :: tminv(*), g2d(*), x(*), y(*)
    complex(c_double_complex), intent(inout) :: tau(*)
    integer(c_int) :: ierr, bg
    bg = 0
    if (bulkgeo) bg = 1
    ierr = skkr_tau2d_g2d(dctx, irel, wgeff, bg, lmax, lkmax, lkmaxp, nintfc, tminv, g2d, x, y, tau, &
         mdimp, d_lmmaxp, d_minprc, d_mprc1, d_mintfc, d_nprc, d_ninprc, real(d_fac, c_double), &
         aimag(d_fac))
    if (ierr /= SKKR_OK) then
      write(*, '(a,i3)') ' tau2d_gpu: error ', ierr
      stop 1
    end if
  end subroutine tau2d_gpu

  subroutine dugo_gpu(lmax, lkmax, tminvb, lkmaxp, g2d, g2di, xy, mdimp, side, bulk, irel, sgfconv)
    integer, intent(in) :: lmax, lkmax, lkmaxp, mdimp, irel
    complex(c_double_complex), intent(in) :: tminvb(*), g2d(*), g2di(*)
    complex(c_double_complex), intent(inout) :: xy(*)
    character(len=1), intent(in) :: side
    logical, intent(in) :: bulk
    logical, intent(out) :: sgfconv
    integer(c_int) :: ierr, conv, its, bk
    bk = 0
    if (bulk) bk = 1
    ierr = skkr_dugo_g2d(dctx, lmax, lkmax, tminvb, lkmaxp, g2d, g2di, xy, mdimp, side, bk, irel, conv, &
         d_lmmaxp, d_minprc, d_nprc, d_ninprc, real(d_fac, c_double), aimag(d_fac), its)
    dropin_last_iters = its
    sgfconv = conv /= 0
    if (ierr /= SKKR_OK .and. ierr /= SKKR_ERR_SGF_DIVERGED) then
      write(*, '(a,i3)') ' dugo_gpu: error ', ierr
      stop 1
    end if
  end subroutine dugo_gpu

end module skkr_dropin
