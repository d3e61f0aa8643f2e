c  Driver of the reference's per-(E,k) subroutines, compiled together with
c  the reference sources where they lie (dugo, surfgfd, packer, tau2d,
c  gjinv, matops of skkr_kpar_v6/main or skkr_dbogo_bsf/main) by
c  oracle/ref/Makefile.  TEST INFRASTRUCTURE: it writes golden vectors.
c
c  It plays the role of the k-loop body of
c  skkr_kpar_v6/rsp-scf/cpacoord_ksave.f:188-391 for a bulk-like stack
c  (bulkgeo=.true.): per k it scatters G(k) into the layer-indexed
c  g2d(lm,lm',jj,icount,icpl) layout that packer.f reads, calls
c  dugo('L'), dugo('R') and tau2d, and accumulates w_k*tau.
c
c  Input (stream, little endian): int32 lk,s,P,nk,itermax,nsave;
c  real*8 errmax; int32 isave(nsave); complex*16 tinv(lk,lk,P*s),
c  tinvl(lk,lk,s), tinvr(lk,lk,s); real*8 w(nk); then per k
c  complex*16 gintra(n,n), ginter(n,n) with n=lk*s (G, not -G).
c  Output: complex*16 taubz(lk,lk,P*s), trk(P*s,nk), tausave(lk,lk,P*s,nsave).
c  Decimation iterations are printed by surfgfd (ichk=1) on stdout.
c  -DGOD (with -DDBOGO): tau2dgod of skkr_dbogo_bsf/clu/tau2d_god.f instead
c  of tau2d, and tausave holds the full tau(lk,lk,P*s,P*s) of the saved k.
c  -DSYM: the BZ sum of an irreducible wedge (intbz=1) as
c  nrsp-scf/cpacoord.f:281-299 forms it: after w(nk) the input holds
c  int32 ng; complex*16 dg(lk,lk,ng), dg1(lk,lk,ng) (sphbas), and every
c  k adds (w/ng) dg tau dg1 of each group element (tripmt, addmatc).
c  -DLAYERED (with -DDBOGO, whose param.h has mprc1 = mprc): a layered
c  stack, tau2d with bulkgeo=.false. (tau2d.f:55-56,90-91).  Per k, after
c  gintra, ginter (the bulk leads, as gg2dbulk of nrsp-scf/cpacoord.f:216-
c  262), P intra blocks gip(n,n) and P-1 couplings gop(n,n) of PL p to PL
c  p+1; plane ipl of g2d holds PL ipl's rows (its couplings to the leads
c  are the bulk ginter).
      program refdrv
      implicit real*8 (a-h,o-z)
      include 'param.h'
#ifdef DBOGO
      parameter(mdim=mdimdbogo, lgd=bogomaxp)
#else
      parameter(mdim=mdimr, lgd=lmmaxp)
#endif
      complex*16 g2d(lgd,lgd,-minprc:minprc,minprc,0:mprc1+1)
#ifdef LAYERED
      complex*16 g2db(lgd,lgd,-minprc:minprc,minprc,0:mprc1+1)
      complex*16, allocatable :: gip(:,:,:),gop(:,:,:)
#endif
      complex*16 xx(mdim,mdim),yy(mdim,mdim)
      complex*16, allocatable :: tinv(:,:,:),tinvl(:,:,:),tinvr(:,:,:)
      complex*16, allocatable :: tau1(:,:,:),taubz(:,:,:)
      complex*16, allocatable :: trk(:,:),tsave(:,:,:,:)
#ifdef GOD
      complex*16, allocatable :: tauf(:,:,:,:),tfsave(:,:,:,:,:)
#endif
      complex*16, allocatable :: gi(:,:),go(:,:)
#ifdef SYM
      complex*16, allocatable :: dg(:,:,:),dg1(:,:,:),tsym(:,:)
      integer ng
#endif
      real*8, allocatable :: w(:)
      integer, allocatable :: isave(:)
      integer lk,ns,np,nk,itmx,nsave
      real*8 errm
      logical sgfconv
      character*512 fin,fout
#ifndef DBOGO
      complex*16 fac
      common/relfac/fac
#endif
      common/test/itest
      common/iterpar/errmax,itermaxl,itermaxr,ichk
      common/lay2d/cvec(mtotal,3),nextra,nbulkl,nbulkr,
     &             nprc,ninprc(0:mprc+1)
c
      call get_command_argument(1,fin)
      call get_command_argument(2,fout)
      open(11,file=fin,access='stream',form='unformatted',status='old')
      read(11) lk,ns,np,nk,itmx,nsave
      read(11) errm
      n=lk*ns
      if(lk.gt.lgd.or.ns.gt.minprc.or.np.gt.mprc.or.n.gt.mdim) then
        write(6,*) 'REFDRV: size exceeds param.h',lk,ns,np,n
        stop 2
      end if
      allocate(isave(max(nsave,1)))
      if(nsave.gt.0) read(11) isave(1:nsave)
      allocate(tinv(lk,lk,mintfc),tinvl(lk,lk,minprc),
     &         tinvr(lk,lk,minprc))
      allocate(tau1(lk,lk,mintfc),taubz(lk,lk,np*ns))
      allocate(trk(np*ns,nk),tsave(lk,lk,np*ns,max(nsave,1)))
      allocate(gi(n,n),go(n,n),w(nk))
#ifdef LAYERED
      allocate(gip(n,n,np),gop(n,n,max(np-1,1)))
#endif
#ifdef GOD
      allocate(tauf(lk,lk,np*ns,np*ns))
      allocate(tfsave(lk,lk,np*ns,np*ns,max(nsave,1)))
#endif
      tinv=(0.d0,0.d0)
      tinvl=(0.d0,0.d0)
      tinvr=(0.d0,0.d0)
      read(11) tinv(:,:,1:np*ns)
      read(11) tinvl(:,:,1:ns)
      read(11) tinvr(:,:,1:ns)
      read(11) w
#ifdef SYM
      read(11) ng
      allocate(dg(lk,lk,ng),dg1(lk,lk,ng),tsym(lk,lk))
      read(11) dg
      read(11) dg1
#endif
c
      itest=0
      errmax=errm
      itermaxl=itmx
      itermaxr=itmx
      ichk=1
      nprc=np
      do i=0,mprc+1
        ninprc(i)=ns
      end do
#ifndef DBOGO
      fac=(1.d0,0.d0)
#endif
      taubz=(0.d0,0.d0)
      lmax=2
c
      do ik=1,nk
        read(11) gi
        read(11) go
        g2d=(0.d0,0.d0)
        do ipl=0,2
          call fillpl(g2d(1,1,-minprc,1,ipl),gi,go,go,lk,ns,n,lgd,
     &                minprc)
        end do
#ifdef LAYERED
        g2db=g2d
        do ipl=1,np
          read(11) gip(:,:,ipl)
        end do
        do ipl=1,np-1
          read(11) gop(:,:,ipl)
        end do
        do ipl=1,np
          if(ipl.eq.1.and.ipl.eq.np) then
            call fillpl(g2d(1,1,-minprc,1,ipl),gip(1,1,ipl),go,go,
     &                  lk,ns,n,lgd,minprc)
          else if(ipl.eq.1) then
            call fillpl(g2d(1,1,-minprc,1,ipl),gip(1,1,ipl),go,
     &                  gop(1,1,ipl),lk,ns,n,lgd,minprc)
          else if(ipl.eq.np) then
            call fillpl(g2d(1,1,-minprc,1,ipl),gip(1,1,ipl),
     &                  gop(1,1,ipl-1),go,lk,ns,n,lgd,minprc)
          else
            call fillpl(g2d(1,1,-minprc,1,ipl),gip(1,1,ipl),
     &                  gop(1,1,ipl-1),gop(1,1,ipl),lk,ns,n,lgd,minprc)
          end if
        end do
#endif
c
#ifdef LAYERED
        call dugo(lmax,lk,tinvl,lk,g2db(1,1,-minprc,1,0),
     &            g2db(1,1,-minprc,1,1),yy,mdim,'L',.true.,0,sgfconv)
        if(.not.sgfconv) stop 3
        call dugo(lmax,lk,tinvr,lk,g2db(1,1,-minprc,1,2),
     &            g2db(1,1,-minprc,1,1),xx,mdim,'R',.true.,0,sgfconv)
        if(.not.sgfconv) stop 3
#else
        call dugo(lmax,lk,tinvl,lk,g2d(1,1,-minprc,1,0),
     &            g2d(1,1,-minprc,1,1),yy,mdim,'L',.true.,0,sgfconv)
        if(.not.sgfconv) stop 3
        call dugo(lmax,lk,tinvr,lk,g2d(1,1,-minprc,1,2),
     &            g2d(1,1,-minprc,1,1),xx,mdim,'R',.true.,0,sgfconv)
        if(.not.sgfconv) stop 3
#endif
#ifdef GOD
        call tau2dgod(0,1.d0,.true.,lmax,lk,lk,np*ns,tinv,g2d,xx,yy,
     &                tauf,mdim)
        do li=1,np*ns
          tau1(:,:,li)=tauf(:,:,li,li)
        end do
        do is=1,nsave
          if(isave(is).eq.ik) tfsave(:,:,:,:,is)=tauf
        end do
#elif defined(LAYERED)
        call tau2d(0,1.d0,.false.,lmax,lk,lk,np*ns,tinv,g2d,xx,yy,
     &             tau1,mdim)
#else
        call tau2d(0,1.d0,.true.,lmax,lk,lk,np*ns,tinv,g2d,xx,yy,
     &             tau1,mdim)
#endif
        do li=1,np*ns
          trk(li,ik)=(0.d0,0.d0)
          do l=1,lk
            trk(li,ik)=trk(li,ik)+tau1(l,l,li)
          end do
#ifdef SYM
          do ig=1,ng
            tsym=tau1(:,:,li)
            call tripmt(dg(1,1,ig),tsym,dg1(1,1,ig),lk,lk,lk)
            call addmatc(taubz(1,1,li),tsym,dcmplx(w(ik)/ng,0.d0),lk,lk)
          end do
#else
          do l=1,lk
          do lp=1,lk
            taubz(l,lp,li)=taubz(l,lp,li)+w(ik)*tau1(l,lp,li)
          end do
          end do
#endif
        end do
        do is=1,nsave
          if(isave(is).eq.ik) tsave(:,:,:,is)=tau1(:,:,1:np*ns)
        end do
      end do
      close(11)
c
      open(12,file=fout,access='stream',form='unformatted',
     &     status='replace')
      write(12) taubz
      write(12) trk
#ifdef GOD
      if(nsave.gt.0) write(12) tfsave(:,:,:,:,1:nsave)
#else
      if(nsave.gt.0) write(12) tsave(:,:,:,1:nsave)
#endif
      close(12)
      end
c
c  Rows of one principal layer in g2d(lm,lm',jj,icount) (packer.f:36-95):
c  row site i couples to PL p-1 through offsets -s..-i (gm, the coupling of
c  PL p-1 to PL p, read transposed and conjugated), within PL p through
c  1-i..s-i (gi) and to PL p+1 through s-i+1..s (gn).  G, not -G.
      subroutine fillpl(g2p,gi,gm,gn,lk,ns,n,lgd,mnp)
      implicit real*8 (a-h,o-z)
      complex*16 g2p(lgd,lgd,-mnp:mnp,mnp)
      complex*16 gi(n,n),gm(n,n),gn(n,n)
      do i=1,ns
        do j=-ns,-i
          jp=ns+i+j
          do ia=1,lk
          do ib=1,lk
            g2p(ia,ib,j,i)=dconjg(gm((jp-1)*lk+ib,(i-1)*lk+ia))
          end do
          end do
        end do
        do j=1-i,ns-i
          jp=i+j
          do ia=1,lk
          do ib=1,lk
            g2p(ia,ib,j,i)=gi((i-1)*lk+ia,(jp-1)*lk+ib)
          end do
          end do
        end do
        do j=ns-i+1,ns
          jp=j+i-ns
          do ia=1,lk
          do ib=1,lk
            g2p(ia,ib,j,i)=gn((i-1)*lk+ia,(jp-1)*lk+ib)
          end do
          end do
        end do
      end do
      end
