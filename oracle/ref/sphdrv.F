c  Driver of the reference's point-group matrices, compiled together with
c  skkr_kpar_v6/nrsp-scf/{sphbas,matrep}.f and main/{matops,zero}.f where
c  they lie, by oracle/ref/Makefile.  TEST INFRASTRUCTURE: it writes golden
c  vectors.
c
c  For each 2D point group of sphbas.f (C1 C2 C3 C4 C6 Cs C2v C3vA C3vB
c  C4v C6v) it calls sphbas (the matrices dg of the group elements in the
c  lm basis, their inverses dg1, the element count ng and the inverse
c  labels invg) as nrsp-scf/cpacoord.f:98 does.
c
c  Output (stream): int32 lmmaxp, melem, ngroups; per group int32 ng,
c  invg(melem); complex*16 dg(lmmaxp,lmmaxp,melem), dg1(same).
c  -DREL: rsp-scf/{sphbas,matrep}.f, the kappa-mu basis (double group),
c  kmymaxp in place of lmmaxp.
      program refsph
      implicit real*8 (a-h,o-z)
      include 'param.h'
      character*4 idgroup
      common/latt2d/idgroup,ilat,a2d,b2d,d2d
#ifdef REL
      parameter(ldg=kmymaxp)
#else
      parameter(ldg=lmmaxp)
#endif
      complex*16 dg(ldg,ldg,melem),dg1(ldg,ldg,melem)
      integer invg(melem)
      character*4 groups(11)
      character*512 fout
      data groups/'C1  ','C2  ','C3  ','C4  ','C6  ','Cs  ','C2v ',
     &            'C3vA','C3vB','C4v ','C6v '/
      call get_command_argument(1,fout)
      open(12,file=fout,access='stream',form='unformatted',
     &     status='replace')
      write(12) ldg,melem,11
      do i=1,11
        idgroup=groups(i)
        dg=(0.d0,0.d0)
        dg1=(0.d0,0.d0)
        invg=0
        call sphbas(dg,dg1,ng,invg)
        write(12) ng
        write(12) invg
        write(12) dg
        write(12) dg1
      end do
      close(12)
      end
