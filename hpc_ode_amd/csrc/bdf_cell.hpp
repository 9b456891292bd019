// The per-cell CVODE BDF integration shared by the wave-per-cell and the block-per-cell
// kernels: the controller the reference configures at src/Integrators/CVodeIntegrator.cpp:98-155
// (BDF, scalar tolerances, dense direct linear solver, analytic Jacobian, maxsteps, one
// CV_NORMAL call from t = 0 to the cell's dt), written once over a thread-group layout.
#pragma once

#include <float.h>

#include "bdf_device.hpp"
#include "kernels.hpp"

namespace hpo {
namespace dev {

// CVODE constants (SUNDIALS 6.x defaults)
constexpr int QMAX = 5;
constexpr double ADDON = 1.0e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0;
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1, ETACF = 0.25;
constexpr double THRESH = 1.5;
constexpr int SMALL_NST = 10, MXNCF = 10, MXNEF = 7, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10;
constexpr int NLS_MAXCOR = 3, MSBP = 20, MSBJ = 51;
constexpr double CRDOWN = 0.3, DGMAX_LSETUP = 0.3, RDIV = 2.0, DGMAX_JBAD = 0.2, NLSCOEF = 0.1;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;
constexpr int MAX_ITERS = 4;
constexpr double UROUND = DBL_EPSILON;

enum { FIRST_CALL = 0, PREV_CONV_FAIL = 1, PREV_ERR_FAIL = 2 };
enum { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };
enum { PH_F0 = 0, PH_HIN = 1, PH_NEWTON = 2, PH_RESTART = 3 };
enum { ACT_EVAL = 0, ACT_BEGIN = 1, ACT_ATTEMPT = 2, ACT_DONE = 3 };

// Word 3 of a launch's work queue: nonzero when the FP64 fallback is armed (read alike by every lane).
HPO_DI bool hpo_fallback_armed(const uint32_t* q) {
  return __builtin_amdgcn_readfirstlane(static_cast<int>(*reinterpret_cast<const volatile uint32_t*>(q + 3))) != 0;
}

// CVODE's weighted RMS norm over the cell's components (one per thread of the cell's group).
template <class Lay>
HPO_DI double wrms(const Lay& Ly, double v, double w, int n) {
  double t = v * w;
  return sqrt(Ly.sum(t * t) / n);
}

// Wave-uniform controller state. It lives in per-wave LDS so that the ~40 scalars
// do not each occupy a VGPR pair next to the register-resident Newton matrix.
struct Scal {
  double tn, h, hprime, hscale, eta, etamax, hu;
  double gamma, gammap, gamrat, rl1, crate, delp, acnrm, saved_tq5, saved_t, gamma_lu;
  double hlb, hub, hg, hin_sign;
  double tau[8], tq[6], l[6];
};
static_assert(sizeof(Scal) % 8 == 0, "Scal layout");
constexpr int kScalDoubles = sizeof(Scal) / 8;

// One cell's CVODE BDF integration on the cell's thread group (a wave or a workgroup), whose
// layout object provides the thread index (tid: the state component the thread owns), the
// group-wide reductions (any, sum, max), the RHS, the Newton-matrix setup and the linear solve.
// Every thread runs the same control flow: all decisions depend only on group-wide values.
template <int NS, bool LU32, bool PROF, class Lay>
__device__ __forceinline__ void integrate_cell(const BdfArgs& A, Lay& Ly, Scal& S, int64_t cell) {
  const int n = NS;
  const int lane = Ly.tid;
  const bool act = lane < n;
  const double P = A.P[cell];
  const double tout = A.dt[cell];
  const double rtol = A.rtol, atol = A.atol;

  double zn[QMAX + 1];
  zn[0] = act ? A.y_in[cell * A.ld + lane] : 0.0;
#pragma unroll
  for (int j = 1; j <= QMAX; ++j) zn[j] = 0.0;
  double ewt = 0.0, acor = 0.0, yv = zn[0];

  if (Ly.wave_lane == 0) {
    S.tn = 0.0;
    S.h = S.hprime = S.hscale = 0.0;
    S.eta = 1.0;
    S.etamax = ETAMX1;
    S.hu = 0.0;
    S.gamma = S.gammap = 0.0;
    S.gamrat = 1.0;
    S.rl1 = 1.0;
    S.crate = 1.0;
    S.delp = S.acnrm = S.saved_tq5 = S.saved_t = S.gamma_lu = 0.0;
    S.hlb = S.hub = S.hg = 0.0;
    S.hin_sign = 1.0;
    for (int j = 0; j < 8; ++j) S.tau[j] = 0.0;
    for (int j = 0; j < 6; ++j) S.tq[j] = S.l[j] = 0.0;
  }
  wave_sync();
  int q = 1, qprime = 1, qwait = 2, L = 2, jcur = 0, convfail = NO_FAILURES;
  int nst = 0, nfe = 0, nje = 0, nsetups = 0, nni = 0, ncfn = 0, netf = 0;
  int nstlp = 0, nstlj = 0, nstloc = 0;
  int ncf = 0, nef = 0, nflag = FIRST_CALL, m = 0, need_setup = 0, jbad = 0;
  int status = HPO_SUCCESS;
  int count1 = 0;
  constexpr bool prof = PROF;   // per-phase cycle counters: a separate kernel instantiation
  long long c_rhs = 0, c_jac = 0, c_lu = 0, c_sol = 0, c_t0 = prof ? clock64() : 0, c_a = 0;

  // error weights for y0 (cvInitialSetup)
  {
    double tol = rtol * fabs(zn[0]) + atol;
    if (Ly.any(act && !(tol > 0.0))) status = HPO_ILL_INPUT;
    ewt = act ? 1.0 / tol : 0.0;
  }

  // All writes to S below happen identically in every lane (same value, same address),
  // so no lane-0 guard is needed; reads see the wave's own writes in program order.
  auto rescale = [&]() {
    double f = S.eta;
#pragma unroll
    for (int j = 1; j <= QMAX; ++j) {
      if (j <= q) {
        zn[j] *= f;
        f *= S.eta;
      }
    }
    S.h = S.hscale * S.eta;
    S.hscale = S.h;
  };
  auto restore = [&]() {
    S.tn = S.saved_t;
#pragma unroll
    for (int k = 1; k <= QMAX; ++k)
#pragma unroll
      for (int j = QMAX; j >= k; --j)
        if (j <= q) zn[j - 1] -= zn[j];
  };
  auto adjust_order = [&](int deltaq) {
    if (q == 2 && deltaq != 1) return;
    if (deltaq == 1) {
      // cvIncreaseBDF
      double alpha0 = -1.0, alpha1 = 1.0, prod = 1.0, xiold = 1.0, hsum = S.hscale;
      for (int i = 0; i <= QMAX; ++i) S.l[i] = 0.0;
      S.l[2] = 1.0;
      for (int j = 1; j < q; ++j) {
        hsum += S.tau[j + 1];
        double xi = hsum / S.hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
        for (int i = j + 2; i >= 2; --i) S.l[i] = S.l[i] * xiold + S.l[i - 1];
        xiold = xi;
      }
      double A1 = (-alpha0 - alpha1) / prod;
      double zl = A1 * zn[QMAX];  // zn[indx_acor], indx_acor = qmax
      put(zn, L, zl);
#pragma unroll
      for (int j = 2; j <= QMAX; ++j)
        if (j <= q) zn[j] += S.l[j] * zl;
    } else if (deltaq == -1) {
      // cvDecreaseBDF
      for (int i = 0; i <= QMAX; ++i) S.l[i] = 0.0;
      S.l[2] = 1.0;
      double hsum = 0.0;
      for (int j = 1; j <= q - 2; ++j) {
        hsum += S.tau[j];
        double xi = hsum / S.hscale;
        for (int i = j + 2; i >= 2; --i) S.l[i] = S.l[i] * xi + S.l[i - 1];
      }
      double zq = pick(zn, q);
#pragma unroll
      for (int j = 2; j < QMAX; ++j)
        if (j < q) zn[j] = -S.l[j] * zq + zn[j];
    }
  };
  // predict + cvSetBDF + cvNls preamble; returns ACT_EVAL
  auto attempt = [&]() -> int {
    S.tn += S.h;
#pragma unroll
    for (int k = 1; k <= QMAX; ++k)
#pragma unroll
      for (int j = QMAX; j >= k; --j)
        if (j <= q) zn[j - 1] += zn[j];
    {
      // cvSetBDF
      // The l polynomial is built in registers (loops unrolled over the order, guarded by
      // the wave-uniform q) and stored once: in LDS every update would be a dependent
      // load-store round trip.
      const double h = S.h;
      double alpha0 = -1.0, alpha0_hat = -1.0, xi_inv = 1.0, xistar_inv = 1.0, hsum = h;
      double l[QMAX + 1] = {1.0, 1.0, 0.0, 0.0, 0.0, 0.0};
      if (q > 1) {
#pragma unroll
        for (int j = 2; j < QMAX; ++j) {
          if (j < q) {
            hsum += S.tau[j - 1];
            xi_inv = h / hsum;
            alpha0 -= 1.0 / j;
#pragma unroll
            for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
          }
        }
        alpha0 -= 1.0 / q;
        xistar_inv = -l[1] - alpha0;
        hsum += S.tau[q - 1];
        xi_inv = h / hsum;
        alpha0_hat = -l[1] - xi_inv;
#pragma unroll
        for (int i = QMAX; i >= 1; --i)
          if (i <= q) l[i] += l[i - 1] * xistar_inv;
      }
#pragma unroll
      for (int i = 0; i <= QMAX; ++i) S.l[i] = l[i];
      // cvSetTqBDF
      double A1 = 1.0 - alpha0_hat + alpha0;
      double A2 = 1.0 + q * A1;
      double lq = pick(l, q);
      S.tq[2] = fabs(A1 / (alpha0 * A2));
      S.tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
      if (qwait == 1) {
        if (q > 1) {
          double C = xistar_inv / lq;
          double A3 = alpha0 + 1.0 / q;
          double A4 = alpha0_hat + xi_inv;
          double Cpinv = (1.0 - A4 + A3) / A3;
          S.tq[1] = fabs(C * Cpinv);
        } else {
          S.tq[1] = 1.0;
        }
        hsum += S.tau[q];
        xi_inv = h / hsum;
        double A5 = alpha0 - (1.0 / (q + 1));
        double A6 = alpha0_hat - xi_inv;
        double Cppinv = (1.0 - A6 + A5) / A2;
        S.tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
      }
      S.tq[4] = NLSCOEF / S.tq[2];
      S.rl1 = 1.0 / l[1];
      S.gamma = h * S.rl1;
      if (nst == 0) S.gammap = S.gamma;
      S.gamrat = (nst > 0) ? S.gamma / S.gammap : 1.0;
    }
    // cvNls preamble
    convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
    need_setup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (nst == 0) || (nst >= nstlp + MSBP) ||
                 (fabs(S.gamrat - 1.0) > DGMAX_LSETUP);
    jbad = 0;
    m = 0;
    acor = 0.0;
    yv = zn[0];
    return ACT_EVAL;
  };
  // top of an internal step (CVode loop body + cvStep entry); returns ACT_ATTEMPT or ACT_DONE
  auto begin = [&]() -> int {
    if (nst > 0) {
      double tol = rtol * fabs(zn[0]) + atol;
      if (Ly.any(act && !(tol > 0.0))) {
        status = HPO_ILL_INPUT;
        yv = zn[0];
        return ACT_DONE;
      }
      ewt = act ? 1.0 / tol : 0.0;
    }
    if (A.maxsteps > 0 && nstloc >= A.maxsteps) {
      status = HPO_TOO_MUCH_WORK;
      yv = zn[0];
      return ACT_DONE;
    }
    if (UROUND * wrms(Ly, zn[0], ewt, n) > 1.0) {
      status = HPO_TOO_MUCH_ACC;
      yv = zn[0];
      return ACT_DONE;
    }
    S.saved_t = S.tn;
    ncf = 0;
    nef = 0;
    nflag = FIRST_CALL;
    if (nst > 0 && S.hprime != S.h) {
      if (qprime != q) {
        adjust_order(qprime - q);
        q = qprime;
        L = q + 1;
        qwait = L;
      }
      rescale();
    }
    return ACT_ATTEMPT;
  };
  // Nonlinear-solver failure (cvHandleNFlag with a recoverable flag)
  auto nls_fail = [&]() -> int {
    ncfn++;
    restore();
    ncf++;
    S.etamax = 1.0;
    if (ncf == MXNCF) {
      status = HPO_CONV_FAILURE;
      yv = zn[0];
      return ACT_DONE;
    }
    S.eta = ETACF;
    nflag = PREV_CONV_FAIL;
    rescale();
    return ACT_ATTEMPT;
  };
  // Newton converged: error test, then complete the step or retry
  auto nls_success = [&]() -> int {
    double dsm = S.acnrm * S.tq[2];
    if (dsm > 1.0) {
      nef++;
      netf++;
      nflag = PREV_ERR_FAIL;
      restore();
      if (nef == MXNEF) {
        status = HPO_ERR_FAILURE;
        yv = zn[0];
        return ACT_DONE;
      }
      S.etamax = 1.0;
      if (nef <= MXNEF1) {
        double eta = 1.0 / (fast_pow(BIAS2 * dsm, 1.0 / L, Ly.exptab) + ADDON);
        eta = fmax(ETAMIN, eta);
        if (nef >= SMALL_NEF) eta = fmin(eta, ETAMXF);
        S.eta = eta;
        rescale();
        return ACT_ATTEMPT;
      }
      if (q > 1) {
        S.eta = ETAMIN;
        adjust_order(-1);
        L = q;
        q--;
        qwait = L;
        rescale();
        return ACT_ATTEMPT;
      }
      S.eta = ETAMIN;
      S.h *= ETAMIN;
      S.hscale = S.h;
      qwait = LONG_WAIT;
      yv = zn[0];
      return ACT_EVAL;  // PH_RESTART: needs f(tn, zn[0])
    }
    // cvCompleteStep
    nst++;
    const double h = S.h;
    S.hu = h;
    {
      // shift the step-size history: all loads before the stores, not one round trip each
      double tau[QMAX];
#pragma unroll
      for (int i = 1; i < QMAX; ++i) tau[i] = S.tau[i];
#pragma unroll
      for (int i = QMAX; i >= 2; --i)
        if (i <= q) S.tau[i] = tau[i - 1];
      if (q == 1 && nst > 1) S.tau[2] = tau[1];
      S.tau[1] = h;
    }
#pragma unroll
    for (int j = 0; j <= QMAX; ++j)
      if (j <= q) zn[j] += S.l[j] * acor;
    qwait--;
    if (qwait == 1 && q != QMAX) {
      zn[QMAX] = acor;
      S.saved_tq5 = S.tq[5];
    }
    // cvPrepareNextStep
    if (S.etamax == 1.0) {
      qwait = qwait > 2 ? qwait : 2;
      qprime = q;
      S.hprime = h;
      S.eta = 1.0;
    } else {
      double etaq = 1.0 / (fast_pow(BIAS2 * dsm, 1.0 / L, Ly.exptab) + ADDON);
      double eta;
      if (qwait != 0) {
        eta = etaq;
        qprime = q;
      } else {
        qwait = 2;
        double etaqm1 = 0.0, etaqp1 = 0.0;
        if (q > 1) {
          double ddn = wrms(Ly, pick(zn, q), ewt, n) * S.tq[1];
          etaqm1 = 1.0 / (fast_pow(BIAS1 * ddn, 1.0 / q, Ly.exptab) + ADDON);
        }
        if (q != QMAX && S.saved_tq5 != 0.0) {
          double rr = h / S.tau[2], pw = 1.0;
          for (int i = 0; i < L; ++i) pw *= rr;
          double cquot = (S.tq[5] / S.saved_tq5) * pw;
          double tempv = -cquot * zn[QMAX] + acor;
          double dup = wrms(Ly, tempv, ewt, n) * S.tq[3];
          etaqp1 = 1.0 / (fast_pow(BIAS3 * dup, 1.0 / (L + 1), Ly.exptab) + ADDON);
        }
        double etam = fmax(etaqm1, fmax(etaq, etaqp1));
        if (etam < THRESH) {
          eta = 1.0;
          qprime = q;
        } else if (etam == etaqm1) {
          eta = etaqm1;
          qprime = q - 1;
        } else if (etam == etaq) {
          eta = etaq;
          qprime = q;
        } else {
          eta = etaqp1;
          qprime = q + 1;
          zn[QMAX] = acor;
        }
      }
      // cvSetEta
      if (eta < THRESH) {
        eta = 1.0;
        S.hprime = h;
      } else {
        eta = fmin(eta, S.etamax);
        S.hprime = h * eta;
      }
      S.eta = eta;
    }
    S.etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    acor *= S.tq[2];
    nstloc++;
    if ((S.tn - tout) * h >= 0.0) {
      // CVodeGetDky(tout, 0)
      double s = (tout - S.tn) / h;
      double cj = 1.0;
      for (int i = 0; i < q; ++i) cj *= s;
      double yo = cj * pick(zn, q);
#pragma unroll
      for (int j = QMAX - 1; j >= 0; --j) {
        if (j < q) {
          double c = 1.0;
          for (int i = 0; i < j; ++i) c *= s;
          yo += c * zn[j];
        }
      }
      yv = yo;
      status = HPO_SUCCESS;
      return ACT_DONE;
    }
    return ACT_BEGIN;
  };

  int phase = PH_F0;
  int action = (status == HPO_SUCCESS) ? ACT_EVAL : ACT_DONE;
  while (action != ACT_DONE) {
    // ---- the single RHS site ----
    if (prof) c_a = clock64();
    double f = Ly.rhs(yv, P);
    if (prof) c_rhs += clock64() - c_a;
    nfe++;
    action = ACT_DONE;
    if (phase == PH_NEWTON) {
      double delta = S.rl1 * zn[1] + acor - S.gamma * f;
      int fail = 0;
      if (need_setup) {
        need_setup = 0;
        if (jbad) convfail = FAIL_BAD_J;
        // cvLsSetup / cvLsLinSys
        const double gamma = S.gamma;
        double dgamma = fabs(gamma / S.gammap - 1.0);
        bool jnew = (nst == 0) || (nst >= nstlj + MSBJ) || (convfail == FAIL_BAD_J && dgamma < DGMAX_JBAD) ||
                    (convfail == FAIL_OTHER);
        long long cj = 0;
        if (prof) c_a = clock64();
        const bool ok = Ly.template setup<PROF>(yv, P, gamma, jnew, cj);
        if (prof) {
          c_jac += cj;
          c_lu += clock64() - c_a - cj;
        }
        if (jnew) {
          nje++;
          nstlj = nst;
        }
        jcur = jnew ? 1 : 0;
        S.gamma_lu = gamma;
        nsetups++;
        S.gamrat = 1.0;
        S.gammap = gamma;
        S.crate = 1.0;
        nstlp = nst;
        if (!ok) fail = 1;  // singular: recoverable failure returned to the step
      }
      if (fail) {
        action = nls_fail();
      } else {
        nni++;
        double b = -delta;
        double x;
        if (prof) c_a = clock64();
        x = Ly.solve(b, S.gamma_lu, A.refine_sweeps);
        if (prof) c_sol += clock64() - c_a;
        if (!act) x = 0.0;
        const double gamrat = S.gamrat;
        if (gamrat != 1.0) x *= 2.0 / (1.0 + gamrat);
        delta = x;
        acor += delta;
        double del = wrms(Ly, delta, ewt, n);
        double crate = S.crate;
        if (m > 0) {
          crate = fmax(CRDOWN * crate, del / S.delp);
          S.crate = crate;
        }
        double dcon = del * fmin(1.0, crate) / S.tq[4];
        if (dcon <= 1.0) {
          S.acnrm = (m == 0) ? del : wrms(Ly, acor, ewt, n);
          jcur = 0;
          action = nls_success();
          if (action == ACT_EVAL) phase = PH_RESTART;
        } else {
          bool conv_fail = (m >= 1 && del > RDIV * S.delp);
          if (!conv_fail) {
            S.delp = del;
            m++;
            conv_fail = (m >= NLS_MAXCOR);
          }
          if (!conv_fail) {
            yv = zn[0] + acor;
            action = ACT_EVAL;  // next Newton iterate
          } else if (!jcur) {
            need_setup = 1;
            jbad = 1;
            m = 0;
            acor = 0.0;
            yv = zn[0];
            action = ACT_EVAL;
          } else {
            action = nls_fail();
          }
        }
      }
    } else if (phase == PH_RESTART) {
      zn[1] = S.h * f;
      action = ACT_ATTEMPT;
    } else if (phase == PH_F0) {
      zn[1] = f;
      // cvHin
      const double tn = S.tn;
      double tdiff = tout - tn;
      if (tdiff == 0.0) {
        status = HPO_TOO_CLOSE;
        action = ACT_DONE;
      } else {
        double sign = tdiff > 0.0 ? 1.0 : -1.0;
        S.hin_sign = sign;
        double tdist = fabs(tdiff);
        double tround = UROUND * fmax(fabs(tn), fabs(tout));
        if (tdist < 2.0 * tround) {
          status = HPO_TOO_CLOSE;
          action = ACT_DONE;
        } else {
          double hlb = HLB_FACTOR * tround;
          double t1 = HUB_FACTOR * fabs(zn[0]) + (rtol * fabs(zn[0]) + atol);
          double hub_inv = Ly.max(act ? fabs(zn[1]) / t1 : 0.0);
          double hub = HUB_FACTOR * tdist;
          if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
          double hg = sqrt(hlb * hub);
          S.hlb = hlb;
          S.hub = hub;
          S.hg = hg;
          if (hub < hlb) {
            S.h = sign * hg;
            S.hscale = S.h;
            S.hprime = S.h;
            zn[1] *= S.h;
            action = ACT_BEGIN;
          } else {
            count1 = 1;
            yv = (hg * sign) * zn[1] + zn[0];
            phase = PH_HIN;
            action = ACT_EVAL;
          }
        }
      }
    } else {  // PH_HIN: cvYddNorm result
      const double hg = S.hg, hub = S.hub, sign = S.hin_sign;
      double hgs = hg * sign;
      double tempv = (f - zn[1]) * (1.0 / hgs);
      double yddnrm = wrms(Ly, tempv, ewt, n);
      double hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
      bool fin = false;
      if (count1 == MAX_ITERS) {
        fin = true;
      } else {
        double hrat = hnew / hg;
        if (hrat > 0.5 && hrat < 2.0) {
          fin = true;
        } else if (count1 > 1 && hrat > 2.0) {
          hnew = hg;
          fin = true;
        }
      }
      if (!fin) {
        S.hg = hnew;
        count1++;
        yv = (hnew * sign) * zn[1] + zn[0];
        action = ACT_EVAL;
      } else {
        double h0 = H_BIAS * hnew;
        if (h0 < S.hlb) h0 = S.hlb;
        if (h0 > hub) h0 = hub;
        S.h = sign * h0;
        S.hscale = S.h;
        S.hprime = S.h;
        zn[1] *= S.h;
        action = ACT_BEGIN;
      }
    }
    // run the step bookkeeping until the next RHS evaluation or the end
    while (action == ACT_BEGIN || action == ACT_ATTEMPT) {
      if (action == ACT_BEGIN) action = begin();
      if (action == ACT_ATTEMPT) {
        action = attempt();
        phase = PH_NEWTON;
      }
    }
  }
  if (act) A.y_out[cell * A.ld + lane] = yv;
  if (lane == 0) {
    A.status[cell] = status;
    if (A.stats) {
      int* st = A.stats + cell * HPO_NSTATS;
      st[0] = nst;
      st[1] = nfe;
      st[2] = nje;
      st[3] = nsetups;
      st[4] = nni;
      st[5] = ncfn;
      st[6] = netf;
      st[7] = q;
    }
    if (prof) {
      int64_t* pr = A.prof + cell * HPO_NPROF;
      pr[0] = c_rhs;
      pr[1] = c_jac;
      pr[2] = c_lu;
      pr[3] = c_sol;
      pr[4] = clock64() - c_t0;
    }
  }
  wave_sync();
}

}  // namespace dev
}  // namespace hpo
