c  Driver of the reference's lm -> kappa-mu transformation of the structure
c  constants: clebsch (skkr_kpar_v6/main/clebsch.f) and relmtrx
c  (main/relmtrx.f), compiled where they lie by oracle/ref/Makefile.
c  TEST INFRASTRUCTURE: it writes golden vectors.
c
c  Input (stream): int32 ncase, lmax; per case complex*16 fac,
c  complex*16 a(nl2,nl2).  Output: per case complex*16 b(2 nl2, 2 nl2).
      program reldrv
      implicit real*8 (a-h,o-z)
      include 'param.h'
      complex*16 a(lmmaxp,lmmaxp),b(kmymaxp,kmymaxp),fac
      integer ncase,lmax
      character*512 fin,fout
      call get_command_argument(1,fin)
      call get_command_argument(2,fout)
      open(11,file=fin,access='stream',form='unformatted',status='old')
      open(12,file=fout,access='stream',form='unformatted',
     &     status='replace')
      read(11) ncase,lmax
      nl2=(lmax+1)**2
      if(nl2.gt.lmmaxp) stop 2
      call clebsch
      do ic=1,ncase
        a=(0.d0,0.d0)
        read(11) fac
        read(11) a(1:nl2,1:nl2)
        call relmtrx(a,b,fac,lmax)
        write(12) b(1:2*nl2,1:2*nl2)
      end do
      close(11)
      close(12)
      end
