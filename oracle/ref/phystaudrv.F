c  Driver of the reference's phystau (screened -> physical tau), compiled
c  together with the reference source where it lies by oracle/ref/Makefile:
c    default : skkr_kpar_v6/nrsp-scf/phystau.f (lm basis, lml from lmfill)
c    -DRSP   : skkr_kpar_v6/rsp-scf/phystau.f  (kappa-mu basis, ldex)
c    -DDBOGO : skkr_dbogo_bsf/rsp-scf/phystau_d.f (electron/hole halves)
c  TEST INFRASTRUCTURE: it writes golden vectors.
c
c  Input (stream): int32 ncase,lmax,idiag,n; per case complex*16
c  tau(n,n), tinv1(n,n), tinv2(n,n), alpha(0:lmax) [, alphah(0:lmax)].
c  Output: per case complex*16 tau(n,n) after phystau.
      program phystaudrv
      implicit real*8 (a-h,o-z)
      include 'param.h'
#if defined(DBOGO)
      parameter(ld=dbogomaxp)
#elif defined(RSP)
      parameter(ld=kmymaxp)
#else
      parameter(ld=lmmaxp)
#endif
      complex*16 tau(ld,ld),t1(ld,ld),t2(ld,ld)
      complex*16 alpha(0:lmaxp),alphah(0:lmaxp)
      integer ncase,lmax,idiag,n
      character*512 fin,fout
c
      call get_command_argument(1,fin)
      call get_command_argument(2,fout)
      open(11,file=fin,access='stream',form='unformatted',status='old')
      open(12,file=fout,access='stream',form='unformatted',
     &     status='replace')
      read(11) ncase,lmax,idiag,n
      if(lmax.gt.lmaxp.or.n.gt.ld) then
        write(6,*) 'PHYSTAUDRV: size exceeds param.h',lmax,n
        stop 2
      end if
#if !defined(DBOGO) && !defined(RSP)
      call lmfill
#endif
      do ic=1,ncase
        call czero(tau,ld*ld)
        call czero(t1,ld*ld)
        call czero(t2,ld*ld)
        read(11) tau(1:n,1:n)
        read(11) t1(1:n,1:n)
        read(11) t2(1:n,1:n)
        read(11) alpha(0:lmax)
#if defined(DBOGO)
        read(11) alphah(0:lmax)
        call phystau(tau,ld,t1,t2,alpha,alphah,lmax,idiag)
#else
        call phystau(tau,t1,t2,alpha,lmax,idiag)
#endif
        write(12) tau(1:n,1:n)
      end do
      close(11)
      close(12)
      end
