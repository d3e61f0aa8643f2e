c  Driver of the reference's 2D k-mesh (kmesh2d) with the lattice set-up it
c  relies on, compiled together with skkr_kpar_v6/main/{primkr2d,rgroup,
c  kmesh2d,speck2,speck,setspk,ibz,fullbz,findangles,vecrot,matops,zero}.f
c  where they lie, by oracle/ref/Makefile.  TEST INFRASTRUCTURE: it writes
c  golden vectors.
c
c  Per case it does what nrsp-scf/readgeom.f:312,564-569 and
c  cpacoord.f do before the k-loop: primkr2d (lattice, reciprocal vectors,
c  phik), rgroup for the input group (a0mat) and for the holohedral group
c  of the lattice (amat, igroup of readgeom.f:51), then kmesh2d.
c
c  Input (stream): int32 ncase; per case int32 ilat, intbz, kunning, iset;
c  real*8 a2d, b2d, d2d, park(2); character*4 idgroup.
c  Output (stream): per case int32 nk; real*8 xk(nk), yk(nk), wk(nk).
      program refkmesh
      implicit real*8 (a-h,o-z)
      include 'param.h'
      parameter(mkd=mkpar)
      character*4 idgroup,iogroup,imgroup,igroup(5),idin
      common/latt2d/idgroup,i2dlat,a2d,beta,delta
      common/brav2d/ ax(2),ay(2)
      common/brill2d/ bx(2),by(2)
      common/cunnbz/ bxc(2),byc(2),phik,iogroup
      common/rrotmat/amat(2,2,melem),a0mat(2,2,melem),ng,ng0
      dimension xk(mkd),yk(mkd),wk(mkd),park(2)
      character*512 fin,fout
      data igroup/'C2','C2v','C2v','C4v','C6v'/
      call get_command_argument(1,fin)
      call get_command_argument(2,fout)
      open(11,file=fin,access='stream',form='unformatted',status='old')
      open(12,file=fout,access='stream',form='unformatted',
     &     status='replace')
      read(11) ncase
      do ic=1,ncase
        read(11) i2dlat,intbz,kun,iset
        read(11) a2d,beta,delta,park
        read(11) idin
        idgroup=idin
        call primkr2d(i2dlat,idgroup,a2d,beta,delta,arot)
        call rgroup(i2dlat,idgroup,ng0,a0mat,melem)
        imgroup=igroup(i2dlat)
        call rgroup(i2dlat,imgroup,ng,amat,melem)
        call kmesh2d(intbz,kun,iset,park,xk,yk,wk,nk,mkd)
        write(12) nk
        write(12) xk(1:nk)
        write(12) yk(1:nk)
        write(12) wk(1:nk)
      end do
      close(11)
      close(12)
      end
