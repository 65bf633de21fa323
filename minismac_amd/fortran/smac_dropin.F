c Drop-in replacements for the reference's lrelax and lusgs (ref/lrelax.f:51,
c ref/lusgs.f:51) that run them on the GPU through libsmac_amd.so.  Link
c these instead of lrelax.o / lusgs.o: the callers (ref/step.f:538,
c ref/step.f:552, ref/tursa.f:252) are unchanged.  Settings come from the
c reference's COMMON blocks (/impli/, /implr/, /bcwake/); a non-zero status
c stops the run as the reference stops on its own configuration errors.
c Preprocess with cpp -P -traditional -DD_PRECISION against the reference's
c include directory, like the reference sources.

      subroutine smac_fill_params(p)
      use smac_amd
#include "common.f"
      type(smac_params) :: p
      p%alpha = alpha
      p%beta = beta
      p%reynum = reynum
      p%dtau = dtau
      p%dt = dt
      p%underr = underr
      p%cdis = cdis
      p%pin = pin
      p%pout = pout
      p%chord = 1.0d0
      p%xmom = 0.25d0
      p%ymom = 0.0d0
      p%impsch = impsch
      p%ivis = ivis
      p%iturb = iturb
      p%iflxo = iflxo
      p%njsp = njsp
      p%nksp = nksp
      p%nbcimp = nbcimp
      p%ntjsp = ntjsp
      p%ntksp = ntksp
      p%precision = SMAC_FP64
      p%line_order = SMAC_ORDER_REFERENCE
      return
      end

c wake-cut pairs (jt, jb) in the order bcimpds visits them (ref/bc.f:571-591)
      subroutine smac_wake_pairs(npair, wake)
#include "common.f"
      __INTEGER npair, wake(2, *)
      npair = 0
      do 20 nireg = 1, nbcreg_wake
         jnew = 0
         do 10 jt = jbcb_t_wake(nireg), jbce_t_wake(nireg)
            jnew = jnew + 1
            npair = npair + 1
            wake(1, npair) = jt
            wake(2, npair) = jbcb_b_wake(nireg) - jnew + 1
10       continue
20    continue
      return
      end

      subroutine lrelax(jmax,kmax,q,dq,s,btc,bjm,bjp,bkm,bkp)
      use iso_c_binding
      use smac_amd
#include "common.f"
      __INTEGER jmax, kmax
      __REAL q(jmax,kmax,3),dq(jmax,kmax,3),s(jmax,kmax,3),
     &  btc(jmax,kmax,3,3),bjm(jmax,kmax,3,3),bjp(jmax,kmax,3,3),
     &  bkm(jmax,kmax,3,3), bkp(jmax,kmax,3,3)
      type(smac_params) :: p
      integer(c_int) :: rc, npair, wake(2, jkmax)
      call smac_fill_params(p)
      call smac_wake_pairs(npair, wake)
      rc = smac_lrelax(p, jmax, kmax, npair, wake, q, dq, s,
     &                 btc, bjm, bjp, bkm, bkp)
      if (rc .ne. SMAC_OK) then
        print *, 'smac_lrelax failed: ', trim(smac_last_error_f())
        stop 'stop: GPU lrelax failed'
      endif
      return
      end

      subroutine lusgs(jmax,kmax,neqs,q,dq,rhs,
     &                 btc,bjm,bjp,bkm,bkp,scr1)
      use iso_c_binding
      use smac_amd
#include "common.f"
      __INTEGER jmax, kmax, neqs
      __REAL q(jmax,kmax,neqs), dq(jmax,kmax,neqs),
     &  rhs(jmax,kmax,neqs),
     &  bjm(jmax,kmax,neqs,neqs), btc(jmax,kmax,neqs,neqs),
     &  bjp(jmax,kmax,neqs,neqs), bkm(jmax,kmax,neqs,neqs),
     &  bkp(jmax,kmax,neqs,neqs), scr1(jmax,kmax,neqs,neqs)
      type(smac_params) :: p
      integer(c_int) :: rc, npair, wake(2, jkmax)
      call smac_fill_params(p)
      call smac_wake_pairs(npair, wake)
      rc = smac_lusgs(p, jmax, kmax, neqs, npair, wake, q, dq, rhs,
     &                btc, bjm, bjp, bkm, bkp, scr1)
      if (rc .ne. SMAC_OK) then
        print *, 'smac_lusgs failed: ', trim(smac_last_error_f())
        stop 'stop: GPU lusgs failed'
      endif
      return
      end
