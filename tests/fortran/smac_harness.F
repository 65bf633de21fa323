c Runs the drop-in lrelax / lusgs of minismac_amd/fortran/smac_dropin.F -- the routines a
c maintainer links in place of ref/lrelax.f and ref/lusgs.f -- on a system read from a
c file, and writes what they return.  It exercises the Fortran side of the boundary on the
c GPU: COMMON (/timm/ /visc/ /airfr/ /impli/ /implr/ /iparmi/ /iparmr/ /mass/) into
c smac_params, /bcwake/ into the wake-cut pairs, the reference's own array layouts.
c
c   smac_harness <in> <out>
c   in  (stream): jmax kmax neqs impsch; iflxo ivis iturb njsp nksp nbcimp ntjsp ntksp;
c                 alpha beta reynum dtau dt underr cdis pin pout; nbcreg_wake and per cut
c                 region jbcb_t_wake jbce_t_wake jbcb_b_wake; q s btc bjm bjp bkm bkp
c   out (stream): dq s (the rhs as the routine leaves it), and scr1 for lusgs
c Preprocess and compile like the reference's sources (oracle/Makefile, dropin-harness).
      program smac_harness
      use iso_c_binding
#include "common.f"
      integer jmx, kmx, neqs, isch, ir, nel, nbl
      double precision, allocatable :: q(:), dq(:), s(:)
      double precision, allocatable :: btc(:), bjm(:), bjp(:)
      double precision, allocatable :: bkm(:), bkp(:), scr1(:)
      character(len=512) :: fin, fout
      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      open(10, file=trim(fin), access='stream', form='unformatted',
     &     status='old')
      read(10) jmx, kmx, neqs, isch
      read(10) iflxo, ivis, iturb, njsp, nksp, nbcimp, ntjsp, ntksp
      read(10) alpha, beta, reynum, dtau, dt, underr, cdis
      read(10) pin, pout
      read(10) nbcreg_wake
      do ir = 1, nbcreg_wake
        read(10) jbcb_t_wake(ir), jbce_t_wake(ir), jbcb_b_wake(ir)
      enddo
      impsch = isch
      nel = jmx * kmx * neqs
      nbl = nel * neqs
      allocate(q(nel), dq(nel), s(nel), btc(nbl), bjm(nbl))
      allocate(bjp(nbl), bkm(nbl), bkp(nbl), scr1(nbl))
      read(10) q, s, btc, bjm, bjp, bkm, bkp
      close(10)
      dq = 0.0d0
      scr1 = 0.0d0
      if (isch .eq. 1) then
        call lrelax(jmx, kmx, q, dq, s, btc, bjm, bjp, bkm, bkp)
      else
        call lusgs(jmx, kmx, neqs, q, dq, s, btc, bjm, bjp, bkm,
     &             bkp, scr1)
      endif
      open(11, file=trim(fout), access='stream', form='unformatted',
     &     status='replace')
      write(11) dq, s
      if (isch .eq. 2) write(11) scr1
      close(11)
      end
