! The k-loop body of cpacoord on the GPU
! (skkr_kpar_v6/nrsp-scf/cpacoord.f:147-317, rsp-scf/cpacoord_ksave.f:188-400):
! the Fortran host keeps its k-mesh arrays xk(ik,1:2), wk(ik), its BZ-summed
! screened tau(lkmaxp,lkmaxp,site) and taukdiag(mkpar,0:lmaxp,mintfc)
! (cpacoord.f:36,43), and one call replaces the loop over k (structure
! constants, dugo 'L'/'R', tau2d, BZ sum, and the per-k l-resolved traces of
! phystau(tau) of cpacoord.f:294-309).  The context must hold the screening
! parameters (skkr_set_screening with nl = lmax+1 trace groups, the lm -> l
! map and alphaintkkr per index).  A rank that owns the k-points with
! modulo(ik-1,nprocs)==myrank passes only those; tau is accumulated (+=), so
! the MPI_Allreduce after the loop stays as it is, and only the rank's rows of
! taukdiag are written.  On a decimation failure it stops like the reference
! (cpacoord.f:233-238).
module kloop_gpu
  use iso_c_binding
  use skkr_gpu
  implicit none
contains
  subroutine cpacoord_kloop_gpu(ctx, ce, nk, xk, wk, mkpar, lkmax, lkmaxp, nintfc, mintfc, &
       nl, lmaxp, tau, taukdiag, nprocs, myrank)
    type(c_ptr), intent(in) :: ctx
    complex(8), intent(in) :: ce
    integer, intent(in) :: nk, mkpar, lkmax, lkmaxp, nintfc, mintfc, nl, lmaxp, nprocs, myrank
    real(8), intent(in) :: xk(mkpar, 2), wk(mkpar)
    complex(8), intent(inout) :: tau(lkmaxp, lkmaxp, mintfc)
    complex(8), intent(inout) :: taukdiag(mkpar, 0:lmaxp, mintfc)
    real(c_double), allocatable, target :: kk(:), ww(:)
    complex(c_double_complex), allocatable, target :: acc(:,:,:), accp(:,:,:), tkd(:,:,:)
    integer(c_int), allocatable, target :: iters(:), status(:)
    integer, allocatable :: ikmap(:)
    type(skkr_stats), target :: st
    integer :: ik, m, nmy, li, l, rc
    if (nl - 1 > lmaxp) stop ' cpacoord_kloop_gpu: nl - 1 > lmaxp'
    nmy = 0
    do ik = 1, nk
       if (modulo(ik - 1, nprocs) == myrank) nmy = nmy + 1
    end do
    allocate(kk(2 * max(nmy, 1)), ww(max(nmy, 1)), ikmap(max(nmy, 1)))
    allocate(acc(lkmax, lkmax, nintfc), accp(lkmax, lkmax, nintfc), tkd(max(nmy, 1), nl, nintfc))
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
    accp = (0.d0, 0.d0)
    st%timing_detail = 0
    rc = skkr_tau_bz_phys(ctx, real(ce, c_double), aimag(ce), int(nmy, c_int), kk, ww, acc, accp, &
         c_loc(tkd), c_loc(iters), c_loc(status), c_loc(st))
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
    ! taukdiag(ik, l, li) = sum over the indices of l of phystau(tau(k))
    do li = 1, nintfc
       do l = 0, nl - 1
          do m = 1, nmy
             taukdiag(ikmap(m), l, li) = tkd(m, l + 1, li)
          end do
       end do
    end do
  end subroutine cpacoord_kloop_gpu
end module kloop_gpu
