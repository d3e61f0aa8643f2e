c  Driver of the reference's CPA solver (tata, cpacor, cpasol), compiled
c  together with the reference source where it lies by oracle/ref/Makefile:
c    default : skkr_kpar_v6/nrsp-scf/fullcpa.f (ndim = lmmaxp)
c    -DRSP   : skkr_kpar_v6/rsp-scf/fullcpa.f  (ndim = kmymaxp)
c  TEST INFRASTRUCTURE: it writes golden vectors.
c
c  Input (stream): int32 ncase, n, nsite; per case real*8 conc(nsite),
c  complex*16 ma(n,n,nsite), mb(n,n,nsite), tauc(n,n,nsite).
c  Output per case: complex*16 mc0(n,n,nsite) (tata: the initial coherent
c  t^-1), then after one cpacor step from mc0: mc(n,n,nsite),
c  taua(n,n,nsite), taub(n,n,nsite), real*8 error.
      program cpadrv
      implicit real*8 (a-h,o-z)
      include 'param.h'
#if defined(RSP)
      parameter(ld=kmymaxp)
#else
      parameter(ld=lmmaxp)
#endif
      complex*16 ma(ld,ld,mintfc),mb(ld,ld,mintfc),mc(ld,ld,mintfc)
      complex*16 tauc(ld,ld,mintfc),taua(ld,ld,mintfc)
      complex*16 taub(ld,ld,mintfc)
      real*8 conc(mintfc)
      logical cpacon,concpa,cpatest
      integer ncase,n,nsite
      character*512 fin,fout
c
      call get_command_argument(1,fin)
      call get_command_argument(2,fout)
      open(11,file=fin,access='stream',form='unformatted',status='old')
      open(12,file=fout,access='stream',form='unformatted',
     &     status='replace')
      read(11) ncase,n,nsite
      if(n.gt.ld.or.nsite.gt.mintfc) then
        write(6,*) 'CPADRV: size exceeds param.h',n,nsite
        stop 2
      end if
      do ic=1,ncase
        call czero(ma,ld*ld*mintfc)
        call czero(mb,ld*ld*mintfc)
        call czero(mc,ld*ld*mintfc)
        call czero(tauc,ld*ld*mintfc)
        read(11) conc(1:nsite)
        do li=1,nsite
          read(11) ma(1:n,1:n,li)
        end do
        do li=1,nsite
          read(11) mb(1:n,1:n,li)
        end do
        do li=1,nsite
          read(11) tauc(1:n,1:n,li)
        end do
        do li=1,nsite
          call tata(conc(li),n,mc(1,1,li),ma(1,1,li),mb(1,1,li))
        end do
        do li=1,nsite
          write(12) mc(1:n,1:n,li)
        end do
        itcpa=1
        itcpam=50
        cpatol=1.d-10
        cpacon=.true.
        concpa=.true.
        cpatest=.false.
        call cpacor(conc,mc,ma,mb,tauc,taua,taub,n,nsite,
     >              itcpa,itcpam,cpatol,cpacon,concpa,cpatest,error)
        do li=1,nsite
          write(12) mc(1:n,1:n,li)
        end do
        do li=1,nsite
          write(12) taua(1:n,1:n,li)
        end do
        do li=1,nsite
          write(12) taub(1:n,1:n,li)
        end do
        write(12) error
      end do
      close(11)
      close(12)
      end
