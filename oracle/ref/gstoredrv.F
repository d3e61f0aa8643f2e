c  Driver of the reference's layer structure-constant producer, compiled
c  together with the reference source where it lies by oracle/ref/Makefile
c  (skkr_kpar_v6/main/{gstore,kamstr,directg,scrstr,scrinv,cerf,incgamma,
c  delkam,ilkam,ylmc,gauntcalc,blm,fac,sphrm,vecsrt2d,indekx,vector2d,
c  latgen2d,genlatt2d,lmfill}.f with skkr_kpar_v6/param.h).
c  TEST INFRASTRUCTURE: it writes golden vectors.
c
c  Input (stream), int32 mode first:
c   mode 1 (special functions): int32 nz, complex*16 z(nz)  -> cerf(z)
c          int32 ng, ng x (int32 n, complex*16 x) -> incgamma gam(0:n)
c          int32 nd, nd x (int32 n, complex*16 x, zz) -> delkam del(0:n)
c          int32 ni, ni x (int32 l, complex*16 alpha, real*8 t,
c                          complex*16 kappa) -> ilkam del(0:l)
c   mode 2 (lattice, kamstr/directg, gstore with bulkgeo):
c          int32 lmax, s, nextra, nsep, nk
c          real*8 eta, a1(2), a2(2); complex*16 e
c          real*8 cvec(3, s+1)   layer positions of one PL and the first
c                                layer of the next (a.u.)
c          complex*16 alphal(0:lmax, s), alphar(0:lmax, s)
c          real*8 sep(3, nsep), real*8 park(2, nk)
c     output: int32 maxr, maxk, maxshr, maxshk; real*8 distr, distk,
c             vol; per sep: complex*16 kamstr alm(nl2,nl2) at park(:,1);
c             complex*16 directg gp(nl2,nl2), gm(nl2,nl2) at park(:,1)
c             (sep(3) > 0 only; zeros otherwise);
c             per k: complex*16 g2d(nl2,nl2,-s:s,s,0:2) of gstore side 'L'
c             and then of side 'R'.
      program gstoredrv
      implicit real*8 (a-h,o-z)
      include 'param.h'
      integer lgaunt
      parameter (lgaunt=(2*lmaxp+1)**2)
      complex*16 z,w,cerf,x,zz,akap,ce,alpha
      complex*16 gam(0:lshape+1),del(0:40)
      complex*16 alm(lmmaxp,lmmaxp),gp(lmmaxp,lmmaxp),gm(lmmaxp,lmmaxp)
      complex*16 g2d(lmmaxp,lmmaxp,-minprc:minprc,minprc,0:mprc1+1)
      real*8 gau(lmmaxp*lmmaxp*lgaunt),sep(0:3),rho(2),park(2)
      real*8 a1in(2),a2in(2)
      real*8, allocatable :: seps(:,:),parks(:,:)
      complex*16 alphal(0:lmaxp,minprc),alphar(0:lmaxp,minprc)
      integer mode
      character*512 fin,fout
      external cerf
c
      common/test/itest
      common/lay2d/cvec(mtotal,3),nextra,nbulkl,nbulkr,
     &             nprc,ninprc(0:mprc+1)
      common/dirvec2d/rvec(2,mdir),distr,maxr
      common/vindr2d/indr(mdir),nshr(mr),maxshr
      common/recvec2d/veck(2,mrec),distk,maxk
      common/vindk2d/indk(mrec),nshk(mk),maxshk
      common/brav2d/ax(2),ay(2)
      common/brill2d/bx(2),by(2)
      common/crystl2d/vol,volbz
      common/scrpar/alphalkkr(0:lmaxp,minprc),alpharkkr(0:lmaxp,minprc),
     &              alphaintkkr(0:lmaxp,mintfc)
      complex*16 alphalkkr,alpharkkr,alphaintkkr
c
      itest=0
      pi=4.d0*datan(1.d0)
      call get_command_argument(1,fin)
      call get_command_argument(2,fout)
      open(11,file=fin,access='stream',form='unformatted',status='old')
      open(12,file=fout,access='stream',form='unformatted',
     &     status='replace')
      read(11) mode
      if(mode.eq.1) then
        read(11) nz
        do i=1,nz
          read(11) z
          w=cerf(z)
          write(12) w
        end do
        read(11) ng
        do i=1,ng
          read(11) n,x
          call incgamma(gam,n,x)
          write(12) gam(0:n)
        end do
        read(11) nd
        do i=1,nd
          read(11) n,x,zz
          call delkam(del,n,x,zz)
          write(12) del(0:n)
        end do
        read(11) ni
        do i=1,ni
          read(11) l,alpha,t,akap
          call ilkam(del,l,alpha,t,akap)
          write(12) del(0:l)
        end do
      else
        read(11) lmax,ns,nextra,nsep,nk
        read(11) eta,a1in,a2in,ce
        nl2=(lmax+1)**2
        if(lmax.gt.lmaxp.or.ns.gt.minprc.or.nextra.gt.mextra) then
          write(6,*) 'GSTOREDRV: size exceeds param.h',lmax,ns,nextra
          stop 2
        end if
        do i=1,ns+1
          read(11) cvec(i,1),cvec(i,2),cvec(i,3)
        end do
        do i=1,ns
          read(11) alphal(0:lmax,i)
        end do
        do i=1,ns
          read(11) alphar(0:lmax,i)
        end do
        allocate(seps(3,max(nsep,1)),parks(2,max(nk,1)))
        read(11) seps(:,1:nsep)
        read(11) parks(:,1:nk)
c       lattice: primkr2d's reciprocal basis, then genlatt2d
        ax=a1in
        ay=a2in
        dnt1=ax(1)*ay(2)-ax(2)*ay(1)
        dnt2=ay(1)*(-ax(2))+ay(2)*ax(1)
        bx(1)=2.d0*pi*ay(2)/dnt1
        bx(2)=2.d0*pi*(-ay(1))/dnt1
        by(1)=2.d0*pi*(-ax(2))/dnt2
        by(2)=2.d0*pi*ax(1)/dnt2
        call genlatt2d
        call lmfill
        write(12) maxr,maxk,maxshr,maxshk
        write(12) distr,distk,vol
c       one PL of ns layers per side, bulk-like: nprc=1 interface PL
        nprc=1
        do i=0,mprc+1
          ninprc(i)=ns
        end do
c       side 'R' reads cvec from minind=(nextra+1)*ns+nintfc: copy the PL
        nintfc=ns
        minind=(nextra+1)*ns+nintfc
        do i=1,ns+1
          do j=1,3
            cvec(minind+i,j)=cvec(i,j)
          end do
        end do
        do i=1,ns
          do l=0,lmax
            alphalkkr(l,i)=alphal(l,i)
            alpharkkr(l,i)=alphar(l,i)
          end do
        end do
c       single kamstr / directg calls at park(:,1)
        call gauntcalc(lmax,gau)
        park(1)=parks(1,1)
        park(2)=parks(2,1)
        do is=1,nsep
          sep(1)=seps(1,is)
          sep(2)=seps(2,is)
          sep(3)=seps(3,is)
          sep(0)=dsqrt(sep(1)**2+sep(2)**2+sep(3)**2)
          rho(1)=sep(1)
          rho(2)=sep(2)
          call vecsrt2d(maxr,mr,rvec,rho,distr,indr,nshr,maxshr)
          call czero(alm,lmmaxp*lmmaxp)
          call kamstr(alm,ce,park(1),park(2),lmax,lmmaxp,maxshk,nshk,
     >                indk,veck,maxshr,nshr,indr,rvec,sep,gau,eta)
          write(12) alm(1:nl2,1:nl2)
          call czero(gp,lmmaxp*lmmaxp)
          call czero(gm,lmmaxp*lmmaxp)
          if(sep(3).gt.0.d0) then
            call directg(park,sep(1),ce,lmax,gp,gm)
          end if
          write(12) gp(1:nl2,1:nl2)
          write(12) gm(1:nl2,1:nl2)
        end do
        do ik=1,nk
          park(1)=parks(1,ik)
          park(2)=parks(2,ik)
          call czero(g2d,lmmaxp*lmmaxp*(2*minprc+1)*minprc*(mprc1+2))
          call gstore(park,ce,eta,.true.,'L',lmax,nintfc,g2d)
          write(12) g2d(1:nl2,1:nl2,-ns:ns,1:ns,0:2)
          call czero(g2d,lmmaxp*lmmaxp*(2*minprc+1)*minprc*(mprc1+2))
          call gstore(park,ce,eta,.true.,'R',lmax,nintfc,g2d)
          write(12) g2d(1:nl2,1:nl2,-ns:ns,1:ns,0:2)
        end do
      end if
      close(11)
      close(12)
      end
