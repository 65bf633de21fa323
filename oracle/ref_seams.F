c Instrumentation seams for the oracle build of the reference
c (TEST INFRASTRUCTURE ONLY).  The recipe in oracle/Makefile compiles
c the unmodified reference with cpp renames applied to one file each:
c   step.f   -Dstep=step_ref      (definition only; smac2d.f calls us)
c   lrelax.f -Dlrelax=lrelax_ref  (step.f calls us)
c   lusgs.f  -Dlusgs=lusgs_ref    (step.f and tursa.f call us)
c   geom.f   -Dforce=force_ref    (smac2d.f/iterout calls us)
c   lreltur.f -Dlreltur=lreltur_ref (tursa.f calls us)
c Each wrapper forwards to the reference routine unchanged and records
c what the reference computed (trace_* in ref_trace.c).  Nothing here
c alters the reference arithmetic.

      subroutine step(jmax,kmax,x,y,rtxy,dj,q,dq,s,vnut,turvar)
#include "common.f"
      __INTEGER jmax, kmax
      __REAL q(jmax,kmax,3),dq(jmax,kmax,3),
     &   x(jmax,kmax), y(jmax,kmax), rtxy(jmax,kmax,2,3),
     &   dj(jmax,kmax), s(jmax,kmax,3),
     &   vnut(jmax,kmax), turvar(jmax,kmax,2)
      __REAL clin
      clin = clift
      call trace_state_in(nt,jmax,kmax,q,turvar)
      call step_ref(jmax,kmax,x,y,rtxy,dj,q,dq,s,vnut,turvar)
      call trace_step(nt,resmax0,resmax,jres,kres,divmax,turres,clin)
      call trace_state_out(nt,jmax,kmax,q,turvar,vnut,dq)
      return
      end

#ifndef SEAMS_NO_SOLVERS
      subroutine lrelax(jmax,kmax,q,dq,s,btc,bjm,bjp,bkm,bkp)
#include "common.f"
      __INTEGER jmax, kmax
      __REAL q(jmax,kmax,3),dq(jmax,kmax,3),s(jmax,kmax,3),
     &  btc(jmax,kmax,3,3),bjm(jmax,kmax,3,3),bjp(jmax,kmax,3,3),
     &  bkm(jmax,kmax,3,3), bkp(jmax,kmax,3,3)
      call trace_norm(nt,jmax,kmax,s)
      call trace_lhs(nt,jmax,kmax,q,s,btc,bjm,bjp,bkm,bkp)
      call lrelax_ref(jmax,kmax,q,dq,s,btc,bjm,bjp,bkm,bkp)
      call trace_solve(nt,jmax,kmax,dq,s)
      return
      end

      subroutine lusgs(jmax,kmax,neqs,q,dq,rhs,
     &                 btc,bjm,bjp,bkm,bkp,scr1)
#include "common.f"
      __INTEGER jmax, kmax, neqs
      __REAL q(jmax,kmax,neqs), dq(jmax,kmax,neqs),
     &  rhs(jmax,kmax,neqs),
     &  bjm(jmax,kmax,neqs,neqs), btc(jmax,kmax,neqs,neqs),
     &  bjp(jmax,kmax,neqs,neqs), bkm(jmax,kmax,neqs,neqs),
     &  bkp(jmax,kmax,neqs,neqs), scr1(jmax,kmax,neqs,neqs)
      if (neqs.eq.3) then
        call trace_norm(nt,jmax,kmax,rhs)
        call trace_lhs(nt,jmax,kmax,q,rhs,btc,bjm,bjp,bkm,bkp)
      endif
      call lusgs_ref(jmax,kmax,neqs,q,dq,rhs,
     &               btc,bjm,bjp,bkm,bkp,scr1)
      if (neqs.eq.3) then
        call trace_solve(nt,jmax,kmax,dq,rhs)
      endif
      return
      end

#endif

      subroutine force(jmax,kmax,x,y,q,rtxy,dj,x0,y0,
     &                 fxp,fyp,fxs,fys,fmom,ioverlap)
#include "common.f"
      __INTEGER jmax, kmax, ioverlap
      __REAL x(jmax,kmax),y(jmax,kmax),q(jmax,kmax,3),
     &  rtxy(jmax,kmax,2,3),dj(jmax,kmax)
      call force_ref(jmax,kmax,x,y,q,rtxy,dj,x0,y0,
     &               fxp,fyp,fxs,fys,fmom,ioverlap)
      call trace_force(nt,alpha,fxp,fyp,fxs,fys,fmom)
      return
      end

      subroutine lreltur(jmax,kmax,turvar,ds,rhs,
     &                   sjm,sjp,skm,skp,stc,nodeidp1)
#include "common.f"
      __INTEGER jmax, kmax, nodeidp1
      __REAL turvar(jmax,kmax), ds(jmax,kmax), rhs(jmax,kmax),
     &     sjm(jmax,kmax), sjp(jmax,kmax), skm(jmax,kmax),
     &     skp(jmax,kmax), stc(jmax,kmax)
      call trace_sa(nt,jmax,kmax,turvar,rhs,sjm,sjp,skm,skp,stc)
      call lreltur_ref(jmax,kmax,turvar,ds,rhs,
     &                 sjm,sjp,skm,skp,stc,nodeidp1)
      call trace_sads(nt,jmax,kmax,ds)
      return
      end
