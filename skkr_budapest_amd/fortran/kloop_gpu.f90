! The k-loop body of cpacoord on the GPU
! (skkr_kpar_v6/rsp-scf/cpacoord_ksave.f:188-400): the Fortran host keeps its
! k-mesh arrays xk(ik,1:2), wk(ik) and its BZ-summed tau(lkmaxp,lkmaxp,site),
! and one call replaces the loop over k (structure constants, dugo 'L'/'R',
! tau2d, BZ sum).  A rank that owns the k-points with
! modulo(ik-1,nprocs)==myrank passes only those; tau is accumulated (+=), so
! the MPI_Allreduce after the loop stays as it is.  On a decimation failure it
! stops like the reference (cpacoord.f:233-238).
module kloop_gpu
  use iso_c_binding
  use skkr_gpu
  implicit none
contains
  subroutine cpacoord_kloop_gpu(ctx, ce, nk, xk, wk, mkpar, lkmax, lkmaxp, nintfc, &
       tau, taukdiag, nprocs, myrank)
    type(c_ptr), intent(in) :: ctx
    complex(8), intent(in) :: ce
    integer, intent(in) :: nk, mkpar, lkmax, lkmaxp, nintfc, nprocs, myrank
    real(8), intent(in) :: xk(mkpar, 2), wk(mkpar)
    complex(8), intent(inout) :: tau(lkmaxp, lkmaxp, nintfc)
    complex(8), intent(out) :: taukdiag(mkpar, nintfc)
    real(c_double), allocatable, target :: kk(:), ww(:)
    complex(c_double_complex), allocatable, target :: acc(:,:,:), trk(:,:)
    integer(c_int), allocatable, target :: iters(:), status(:)
    integer, allocatable :: ikmap(:)
    type(skkr_stats), target :: st
    integer :: ik, m, nmy, li, rc
    nmy = 0
    do ik = 1, nk
       if (modulo(ik - 1, nprocs) == myrank) nmy = nmy + 1
    end do
    allocate(kk(2 * max(nmy, 1)), ww(max(nmy, 1)), ikmap(max(nmy, 1)))
    allocate(acc(lkmax, lkmax, nintfc), trk(nintfc, max(nmy, 1)))
    allocate(iters(2 * max(nmy, 1)), status(max(nmy, 1)))
    m = 0
    do ik = 1, nk
       if (modulo(ik - 1, nprocs) == myrank) then
          m = m + 1
          kk(2 * m - 1) = xk(ik, 1)
          kk(2 * m) = xk(ik, 2)
          ww(m) = wk(ik)
          ikmap(m) = ik
       end if
    end do
    acc = (0.d0, 0.d0)
    st%timing_detail = 0
    rc = skkr_tau_bz(ctx, real(ce, c_double), aimag(ce), int(nmy, c_int), kk, ww, acc, &
         c_loc(trk), c_loc(iters), c_loc(status), c_loc(st))
    if (rc /= SKKR_OK) then
       do m = 1, nmy
          if (status(m) /= SKKR_OK) then
             write(6, '('' SGF/inversion failure at k='',2f15.6,'' code'',i3)') &
                  kk(2 * m - 1), kk(2 * m), status(m)
             exit
          end if
       end do
       stop ' stopping due to SGF'
    end if
    do li = 1, nintfc
       tau(1:lkmax, 1:lkmax, li) = tau(1:lkmax, 1:lkmax, li) + acc(:, :, li)
    end do
    taukdiag = (0.d0, 0.d0)
    do m = 1, nmy
       taukdiag(ikmap(m), :) = trk(:, m)
    end do
  end subroutine cpacoord_kloop_gpu
end module kloop_gpu
