! ISO_C_BINDING interface to libskkr_gpu.so (include/skkr_gpu.h).
! complex(c_double_complex) arrays are passed as the complex*16 arrays of the
! reference (interleaved re,im; column-major).
module skkr_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SKKR_OK = 0, SKKR_ERR_ARG = 1, SKKR_ERR_HIP = 2, &
       SKKR_ERR_NOMEM = 3, SKKR_ERR_SINGULAR = 4, SKKR_ERR_SGF_DIVERGED = 5, &
       SKKR_ERR_SGF_ITERMAX = 6, SKKR_ERR_UNSUPPORTED = 7, SKKR_ERR_STR_NOCONV = 8, &
       SKKR_ERR_TOO_LARGE = 9
  integer(c_int), parameter :: SKKR_FP64 = 0, SKKR_MIXED_FP32 = 1, SKKR_MIXED_BF16 = 2, SKKR_MIXED_INT8 = 3

  type, bind(C) :: skkr_config
     integer(c_int) :: lk, sites_per_pl, n_pl
     real(c_double) :: errmax
     integer(c_int) :: itermax_l, itermax_r, precision
     real(c_double) :: fp64_dec_below_imE
     integer(c_int) :: device, batch
  end type skkr_config

  type, bind(C) :: skkr_stats
     real(c_double) :: ms_total, ms_fourier, ms_decimation, ms_sweeps, ms_accumulate, ms_refine
     real(c_double) :: mean_iter_l, mean_iter_r
     integer(c_int) :: max_iter, n_batches
     integer(c_long_long) :: n_launches
     integer(c_int) :: timing_detail
     integer(c_long_long) :: dom_launches
     real(c_double) :: dom_ms, dom_flops
     character(kind=c_char) :: dom_name(32)
     real(c_double) :: dom_bytes, dom_flops_exec
     integer(c_long_long) :: inv_launches
     real(c_double) :: inv_ms, inv_flops, inv_bytes
     integer(c_int) :: redo_batches
  end type skkr_stats

  ! layer structure-constant producer (gstore.f with bulkgeo)
  type, bind(C) :: skkr_str_config
     integer(c_int) :: lmax, s, nextra
     real(c_double) :: eta
     real(c_double) :: a1(2), a2(2)
     integer(c_int) :: device, batch
  end type skkr_str_config

  interface
     integer(c_int) function skkr_create(ctx, cfg) bind(C, name="skkr_create")
       import :: c_int, c_ptr, skkr_config
       type(c_ptr), intent(out) :: ctx
       type(skkr_config), intent(in) :: cfg
     end function skkr_create

     subroutine skkr_destroy(ctx) bind(C, name="skkr_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine skkr_destroy

     integer(c_int) function skkr_set_structure(ctx, n_R, R, gr_intra, gr_inter) &
          bind(C, name="skkr_set_structure")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_R
       real(c_double), intent(in) :: R(*)
       complex(c_double_complex), intent(in) :: gr_intra(*), gr_inter(*)
     end function skkr_set_structure

     integer(c_int) function skkr_set_tinv(ctx, tinv_int, tinv_l, tinv_r) &
          bind(C, name="skkr_set_tinv")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: tinv_int(*), tinv_l(*), tinv_r(*)
     end function skkr_set_tinv

     integer(c_int) function skkr_tau_bz(ctx, e_re, e_im, nk, k, w, tau_bz, tr_tau_k, &
          iters_k, status_k, stats) bind(C, name="skkr_tau_bz")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       real(c_double), value :: e_re, e_im
       integer(c_int), value :: nk
       real(c_double), intent(in) :: k(*), w(*)
       complex(c_double_complex), intent(inout) :: tau_bz(*)
       type(c_ptr), value :: tr_tau_k, iters_k, status_k, stats
     end function skkr_tau_bz

     ! screening alpha_l(site) expanded per basis index (lml / ldex), phystau.f
     integer(c_int) function skkr_set_screening(ctx, nl, l_of_index, a) &
          bind(C, name="skkr_set_screening")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: nl
       integer(c_int), intent(in) :: l_of_index(*)
       complex(c_double_complex), intent(in) :: a(*)
     end function skkr_set_screening

     ! k-loop plus physical tau: tau_bz_phys += phystau(BZ sum), taukdiag(ik,l,li)
     integer(c_int) function skkr_tau_bz_phys(ctx, e_re, e_im, nk, k, w, tau_bz, &
          tau_bz_phys, taukdiag, iters_k, status_k, stats) bind(C, name="skkr_tau_bz_phys")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       real(c_double), value :: e_re, e_im
       integer(c_int), value :: nk
       real(c_double), intent(in) :: k(*), w(*)
       complex(c_double_complex), intent(inout) :: tau_bz(*), tau_bz_phys(*)
       type(c_ptr), value :: taukdiag, iters_k, status_k, stats
     end function skkr_tau_bz_phys

     ! producer of the screened layer structure constants (gstore.f:6)
     integer(c_int) function skkr_str_create(sc, cfg, cvec) bind(C, name="skkr_str_create")
       import :: c_int, c_ptr, c_double, skkr_str_config
       type(c_ptr), intent(out) :: sc
       type(skkr_str_config), intent(in) :: cfg
       real(c_double), intent(in) :: cvec(*)
     end function skkr_str_create

     subroutine skkr_str_destroy(sc) bind(C, name="skkr_str_destroy")
       import :: c_ptr
       type(c_ptr), value :: sc
     end subroutine skkr_str_destroy

     ! g2d(nl2, nl2, -s:s, s, 0:2) per k, as gstore returns it
     integer(c_int) function skkr_gstore_batch(sc, e_re, e_im, alpha, nk, k, g2d, status_k) &
          bind(C, name="skkr_gstore_batch")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: sc
       real(c_double), value :: e_re, e_im
       complex(c_double_complex), intent(in) :: alpha(*)
       integer(c_int), value :: nk
       real(c_double), intent(in) :: k(*)
       complex(c_double_complex), intent(out) :: g2d(*)
       type(c_ptr), value :: status_k
     end function skkr_gstore_batch

     ! take the k-loop's structure constants from the producer
     integer(c_int) function skkr_set_structure_kkr(ctx, sc, alpha) &
          bind(C, name="skkr_set_structure_kkr")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx, sc
       complex(c_double_complex), intent(in) :: alpha(*)
     end function skkr_set_structure_kkr

     ! relativistic k-loop on the producer (rel 1: psq = E + E^2/c^2, 2: E)
     integer(c_int) function skkr_set_structure_kkr_rel(ctx, sc, alpha, rel) &
          bind(C, name="skkr_set_structure_kkr_rel")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx, sc
       complex(c_double_complex), intent(in) :: alpha(*)
       integer(c_int), value :: rel
     end function skkr_set_structure_kkr_rel

     ! CPA loop of one energy (cpacoord.f:107-345, fullcpa.f)
     integer(c_int) function skkr_cpa_bz(ctx, e_re, e_im, nk, k, w, conc, tinv_a, tinv_b, &
          tinv_c, init_ata, itcpam, cpatol, tau_c, tau_a, tau_b, itcpa, cpaerr, iters_k, &
          status_k, stats) bind(C, name="skkr_cpa_bz")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       real(c_double), value :: e_re, e_im
       integer(c_int), value :: nk
       real(c_double), intent(in) :: k(*), w(*), conc(*)
       complex(c_double_complex), intent(in) :: tinv_a(*), tinv_b(*)
       complex(c_double_complex), intent(inout) :: tinv_c(*)
       integer(c_int), value :: init_ata, itcpam
       real(c_double), value :: cpatol
       complex(c_double_complex), intent(out) :: tau_c(*), tau_a(*), tau_b(*)
       integer(c_int), intent(out) :: itcpa
       real(c_double), intent(out) :: cpaerr
       type(c_ptr), value :: iters_k, status_k, stats
     end function skkr_cpa_bz

     ! off-diagonal tau of interface-site pairs (clu/cpacoord_save.f:432-478)
     integer(c_int) function skkr_set_pairs(ctx, npair, pairs, a1, a2) bind(C, name="skkr_set_pairs")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: npair
       integer(c_int), intent(in) :: pairs(*)
       real(c_double), intent(in) :: a1(2), a2(2)
     end function skkr_set_pairs

     integer(c_int) function skkr_tau_bz_pairs(ctx, e_re, e_im, nk, k, w, tau_bz, tau_ij, tau_ji, &
          iters_k, status_k, stats) bind(C, name="skkr_tau_bz_pairs")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       real(c_double), value :: e_re, e_im
       integer(c_int), value :: nk
       real(c_double), intent(in) :: k(*), w(*)
       complex(c_double_complex), intent(inout) :: tau_bz(*), tau_ij(*), tau_ji(*)
       type(c_ptr), value :: iters_k, status_k, stats
     end function skkr_tau_bz_pairs

     ! tau2d (main/tau2d.f:2) for nb k-points: bulkgeo (one set of packer
     ! blocks m00(n,n,nb)) and layered (m00(n,n,P,nb), one set per PL,
     ! tau2d.f:55-56,90-91); tau(lk,lk,P*s,nb) = w * site-diagonal blocks
     integer(c_int) function skkr_tau2d_batch(ctx, nb, m00, m01, m10, x, y, w, tau, status) &
          bind(C, name="skkr_tau2d_batch")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: nb
       complex(c_double_complex), intent(in) :: m00(*), m01(*), m10(*), x(*), y(*)
       real(c_double), value :: w
       complex(c_double_complex), intent(out) :: tau(*)
       integer(c_int), intent(out) :: status(*)
     end function skkr_tau2d_batch

     integer(c_int) function skkr_tau2d_layered_batch(ctx, nb, m00, m01, m10, x, y, w, tau, &
          status) bind(C, name="skkr_tau2d_layered_batch")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: nb
       complex(c_double_complex), intent(in) :: m00(*), m01(*), m10(*), x(*), y(*)
       real(c_double), value :: w
       complex(c_double_complex), intent(out) :: tau(*)
       integer(c_int), intent(out) :: status(*)
     end function skkr_tau2d_layered_batch

     ! tau2d (main/tau2d.f:2-3) with the reference's own arrays, one k-point;
     ! then param.h's lmmaxp, minprc, mprc1, mintfc, common /lay2d/'s nprc,
     ! ninprc(0:nprc+1) and common /relfac/'s fac
     integer(c_int) function skkr_tau2d_g2d(ctx, irel, wgeff, bulkgeo, lmax, lkmax, lkmaxp, nintfc, &
          tminv, g2d, x, y, tau, mdimp, lmmaxp, minprc, mprc1, mintfc, nprc, ninprc, fac_re, fac_im) &
          bind(C, name="skkr_tau2d_g2d")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: irel, bulkgeo, lmax, lkmax, lkmaxp, nintfc, mdimp
       real(c_double), value :: wgeff
       complex(c_double_complex), intent(in) :: tminv(*), g2d(*), x(*), y(*)
       complex(c_double_complex), intent(inout) :: tau(*)
       integer(c_int), value :: lmmaxp, minprc, mprc1, mintfc, nprc
       integer(c_int), intent(in) :: ninprc(*)
       real(c_double), value :: fac_re, fac_im
     end function skkr_tau2d_g2d

     ! dugo (main/dugo.f:9-10) with the reference's own arrays, one k-point;
     ! then lmmaxp, minprc, nprc, ninprc(0:nprc+1), fac, iters (out)
     integer(c_int) function skkr_dugo_g2d(ctx, lmax, lkmax, tminvb, lkmaxp, g2d, g2di, xy, mdimp, &
          side, bulk, irel, sgfconv, lmmaxp, minprc, nprc, ninprc, fac_re, fac_im, iters) &
          bind(C, name="skkr_dugo_g2d")
       import :: c_int, c_ptr, c_double, c_double_complex, c_char
       type(c_ptr), value :: ctx
       integer(c_int), value :: lmax, lkmax, lkmaxp, mdimp, bulk, irel
       complex(c_double_complex), intent(in) :: tminvb(*), g2d(*), g2di(*)
       complex(c_double_complex), intent(inout) :: xy(*)
       character(kind=c_char), value :: side
       integer(c_int), intent(out) :: sgfconv
       integer(c_int), value :: lmmaxp, minprc, nprc
       integer(c_int), intent(in) :: ninprc(*)
       real(c_double), value :: fac_re, fac_im
       integer(c_int), intent(out) :: iters
     end function skkr_dugo_g2d

     ! point group of an irreducible-wedge mesh (sphbas dg, dg1; cpacoord.f:294-299)
     integer(c_int) function skkr_set_point_group(ctx, ng, dg, dg1) bind(C, name="skkr_set_point_group")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: ng
       complex(c_double_complex), intent(in) :: dg(*), dg1(*)
     end function skkr_set_point_group
  end interface
end module skkr_gpu
