// Persistent per-cell BDF kernel for gfx950.
//
// Each wave64 pulls cells from a device work queue and integrates one cell at a
// time from t = 0 to its dt with CVODE's variable-order BDF (the solver the reference
// configures at src/Integrators/CVodeIntegrator.cpp:98-155: BDF, scalar tolerances,
// dense direct linear solver, analytic Jacobian, maxsteps, one CV_NORMAL call).
// The whole integration is resident: Nordsieck history, error weights and the
// Newton matrix row of lane l stay in registers; the RHS is evaluated at one code
// site driven by a small state machine so the kernel is not replicated per caller.
#include <hip/hip_runtime.h>
#include <float.h>

#include <mutex>

#include "bdf_device.hpp"
#include "kernels.hpp"

// Occupancy targets of the BDF kernel (waves per SIMD). At 1 a wave may use the whole
// 512-entry unified VGPR/AGPR file: the FP64 path keeps its 106-register Newton-matrix row,
// the Nordsieck history and the RHS working set there without scratch spills. The FP32
// path holds a 53-register row and fits 256 registers, so it runs two waves per SIMD (a few
// dozen bytes per lane spill, where the RHS and solve working sets peak together).
#ifndef HPO_WAVES_PER_EU_FP64
#define HPO_WAVES_PER_EU_FP64 1
#endif
// Most loads of the refinement's FP64 Jacobian row issued together. Each group costs one
// round trip to the (cache-missing) Jacobian, so fewer, larger groups wait less, at two
// registers per load in flight. GRI's 53 columns as two groups of 27 measured fastest on
// 262,144 cells (1,262 ms against 1,292 ms for seven groups of 8); one group of 53 spills.
#ifndef HPO_JX_GROUP
#define HPO_JX_GROUP 27
#endif
#ifndef HPO_WAVES_PER_EU_FP32
#define HPO_WAVES_PER_EU_FP32 2
#endif

namespace hpo {
namespace dev {

// The refinement's J x product reads its NS columns in jx_groups() equal groups of
// jx_group() loads, NS rounded up to jx_cols(). The padding columns lie inside the wave's
// 64 x 64 Jacobian block, after the NS x NS entries the Jacobian writes, and are zeroed at
// kernel entry.
template <int NS>
constexpr int jx_groups() {
  return (NS + HPO_JX_GROUP - 1) / HPO_JX_GROUP;
}
template <int NS>
constexpr int jx_group() {
  return (NS + jx_groups<NS>() - 1) / jx_groups<NS>();
}
template <int NS>
constexpr int jx_cols() {
  return jx_groups<NS>() * jx_group<NS>();
}

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

template <int NS, bool LU32>
struct LuStore;
template <int NS>
struct LuStore<NS, false> {
  LuReg<NS, double> L;
};
template <int NS>
struct LuStore<NS, true> {
  LuReg<NS, float> L;
};

HPO_DI double wrms(double v, double w, int n) {
  double t = v * w;
  return sqrt(wsum(t * t) / n);
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

template <int NS, bool LU32, bool PROF, bool INV>
__device__ void integrate_cell(const BdfArgs& A, const Tabs& t, const WaveLds& w, const WaveScratch& g, Scal& S,
                               int64_t cell, int lane) {
  const int n = NS;
  const bool act = lane < n;
  const double P = A.P[cell];
  const double tout = A.dt[cell];
  const double rtol = A.rtol, atol = A.atol;

  double zn[QMAX + 1];
  zn[0] = act ? A.y_in[cell * A.ld + lane] : 0.0;
#pragma unroll
  for (int j = 1; j <= QMAX; ++j) zn[j] = 0.0;
  double ewt = 0.0, acor = 0.0, yv = zn[0];
  LuStore<NS, LU32> lu;

  if (lane == 0) {
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
    if (__ballot(act && !(tol > 0.0)) != 0ull) status = HPO_ILL_INPUT;
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
      if (__ballot(act && !(tol > 0.0)) != 0ull) {
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
    if (UROUND * wrms(zn[0], ewt, n) > 1.0) {
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
        double eta = 1.0 / (fast_pow(BIAS2 * dsm, 1.0 / L, t.exptab) + ADDON);
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
      double etaq = 1.0 / (fast_pow(BIAS2 * dsm, 1.0 / L, t.exptab) + ADDON);
      double eta;
      if (qwait != 0) {
        eta = etaq;
        qprime = q;
      } else {
        qwait = 2;
        double etaqm1 = 0.0, etaqp1 = 0.0;
        if (q > 1) {
          double ddn = wrms(pick(zn, q), ewt, n) * S.tq[1];
          etaqm1 = 1.0 / (fast_pow(BIAS1 * ddn, 1.0 / q, t.exptab) + ADDON);
        }
        if (q != QMAX && S.saved_tq5 != 0.0) {
          double rr = h / S.tau[2], pw = 1.0;
          for (int i = 0; i < L; ++i) pw *= rr;
          double cquot = (S.tq[5] / S.saved_tq5) * pw;
          double tempv = -cquot * zn[QMAX] + acor;
          double dup = wrms(tempv, ewt, n) * S.tq[3];
          etaqp1 = 1.0 / (fast_pow(BIAS3 * dup, 1.0 / (L + 1), t.exptab) + ADDON);
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
    double f = rhs(t, w, yv, P, lane, n);
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
        const int ol = opaque(lane);
        // J's columns are packed at the state length: lanes >= NS read entries of other rows
        // (or never-written scratch, still inside the 64 x 64 J block). Their Newton-matrix rows
        // are never pivot candidates and are zeroed when the factors are permuted, and the
        // refinement's products on those lanes are discarded, so nothing they hold reaches the
        // live lanes.
        const double* Jrow = g.J + ol;
        double dgamma = fabs(gamma / S.gammap - 1.0);
        bool jnew = (nst == 0) || (nst >= nstlj + MSBJ) || (convfail == FAIL_BAD_J && dgamma < DGMAX_JBAD) ||
                    (convfail == FAIL_OTHER);
        if constexpr (LU32) {
          if (jnew) {
            if (prof) c_a = clock64();
            jacobian<NS>(t, w, g, yv, P, lane);
            if (prof) c_jac += clock64() - c_a;
            nje++;
            nstlj = nst;
          }
          if (prof) c_a = clock64();
#pragma unroll
          for (int j = 0; j < NS; ++j) {
            double Jv = Jrow[j * NS];
            lu.L.a[j] = static_cast<float>(-gamma * Jv + ((j == ol) ? 1.0 : 0.0));
          }
        } else {
          if (jnew) {
            if (prof) c_a = clock64();
            jacobian<NS>(t, w, g, yv, P, lane);
            if (prof) c_jac += clock64() - c_a;
            nje++;
            nstlj = nst;
          }
          if (prof) c_a = clock64();
#pragma unroll
          for (int j = 0; j < NS; ++j) lu.L.a[j] = -gamma * Jrow[j * NS] + ((j == ol) ? 1.0 : 0.0);
        }
        jcur = jnew ? 1 : 0;
        bool ok;
        if constexpr (INV)
          ok = gj_invert<NS>(lu.L, lane);
        else
          ok = lu_factor<NS>(lu.L, lane);
        if (prof) c_lu += clock64() - c_a;
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
        if constexpr (LU32) {
          const double* Jrow = g.J + opaque(lane);
          float* vbuf = reinterpret_cast<float*>(w.cp);
          float x32 = solve32<NS, INV>(lu.L, static_cast<float>(b), lane, vbuf);
          x = static_cast<double>(x32);
          const double gl = S.gamma_lu;
          for (int sw = 0; sw < A.refine_sweeps; ++sw) {
            // r = b - (I - gamma J) x in FP64, with the FP64 Jacobian row kept in this
            // wave's scratch (cache-resident) rather than in registers. x is broadcast through
            // per-wave LDS (the enthalpy slots, which only the Jacobian uses, after writing them)
            wave_sync();
            w.hRT[lane] = x;
            wave_sync();
            // Whole groups of JG loads are issued before their products, one round trip per
            // group. Columns NS..JN-1 are zero padding (set at kernel entry) and x is zero on
            // lanes >= NS, so the padded products add nothing.
            constexpr int JG = jx_group<NS>();
            double Jx = 0.0;
#pragma unroll 1
            for (int j0 = 0; j0 < jx_cols<NS>(); j0 += JG) {
              double v[JG];
#pragma unroll
              for (int k = 0; k < JG; ++k) v[k] = Jrow[(j0 + k) * NS];
#pragma unroll
              for (int k = 0; k < JG; ++k) Jx = fma(v[k], w.hRT[j0 + k], Jx);
            }
            double r = b - x + gl * Jx;
            float d32 = solve32<NS, INV>(lu.L, static_cast<float>(r), lane, vbuf);
            x += static_cast<double>(d32);
          }
        } else {
          x = lu_solve<NS>(lu.L, b, lane);
        }
        if (prof) c_sol += clock64() - c_a;
        if (!act) x = 0.0;
        const double gamrat = S.gamrat;
        if (gamrat != 1.0) x *= 2.0 / (1.0 + gamrat);
        delta = x;
        acor += delta;
        double del = wrms(delta, ewt, n);
        double crate = S.crate;
        if (m > 0) {
          crate = fmax(CRDOWN * crate, del / S.delp);
          S.crate = crate;
        }
        double dcon = del * fmin(1.0, crate) / S.tq[4];
        if (dcon <= 1.0) {
          S.acnrm = (m == 0) ? del : wrms(acor, ewt, n);
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
          double hub_inv = wmax(act ? fabs(zn[1]) / t1 : 0.0);
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
      double yddnrm = wrms(tempv, ewt, n);
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

// Small mechanisms (at most 16 state variables: a 16-register Newton row, one or two reaction
// passes) fit several waves per SIMD, which hides the latency one wave exposes. On 262,144
// H2/air cells: 918 ms FP32 / 915 ms FP64 at two waves, 861 / 903 ms at three (168 registers,
// a few hundred bytes of spill per lane) and 990 / 1,170 ms at four. Two is the default: one
// three-wave FP64 build returned convergence failures for every cell after an unrelated,
// never-taken branch was added to the cell loop, and no cause has been shown for it.
#ifndef HPO_WAVES_PER_EU_SMALL
#define HPO_WAVES_PER_EU_SMALL 2
#endif
template <int NS, bool LU32>
constexpr int bdf_waves_per_eu() {
  return NS <= 16 ? HPO_WAVES_PER_EU_SMALL : (LU32 ? HPO_WAVES_PER_EU_FP32 : HPO_WAVES_PER_EU_FP64);
}

template <int NS, bool LU32, bool PROF, bool INV>
__global__ __launch_bounds__(256) __attribute__((
    amdgpu_waves_per_eu(bdf_waves_per_eu<NS, LU32>(), bdf_waves_per_eu<NS, LU32>()))) void bdf_kernel(BdfArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const DevMechLayout& Lay = A.lay;
  for (int i = threadIdx.x * 16; i < Lay.lds_bytes; i += blockDim.x * 16)
    *reinterpret_cast<uint4*>(smem + i) = *reinterpret_cast<const uint4*>(A.mech + i);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  Tabs t = make_tabs(smem, A.mech, Lay);
  const int wl_doubles = wave_lds_doubles(Lay.R, Lay.n_partials) + kScalDoubles;
  double* wbase = reinterpret_cast<double*>(smem + Lay.lds_bytes) + wave * wl_doubles;
  WaveLds w = make_wave_lds(wbase, Lay.R, Lay.n_partials);
  Scal& S = *reinterpret_cast<Scal*>(wbase + wave_lds_doubles(Lay.R, Lay.n_partials));
  const int64_t slot = static_cast<int64_t>(blockIdx.x) * wpb + wave;
  WaveScratch g = make_wave_scratch(A.scratch + slot * A.scratch_stride, Lay);
  if (lane < 16) g.coefbuf[lane] = t.coef[lane];
  static_assert((jx_cols<NS>() - 1) * NS + 64 <= 64 * 64, "padded J x columns leave the Jacobian block");
  for (int i = NS * NS + lane; i < (jx_cols<NS>() - 1) * NS + 64; i += 64) g.J[i] = 0.0;
  wave_sync();
  for (;;) {
    unsigned int c = 0;
    if (lane == 0) c = atomicAdd(A.queue, 1u);
    c = __builtin_amdgcn_readfirstlane(c);
    if (static_cast<int64_t>(c) >= A.ncells) break;
    int64_t cell = A.order ? static_cast<int64_t>(A.order[c]) : static_cast<int64_t>(c);
    integrate_cell<NS, LU32, PROF, INV>(A, t, w, g, S, cell, lane);
  }
}

// ---- component kernels used by the parity tests ----
template <int NS, int MODE>
__global__ __launch_bounds__(256) void eval_kernel(EvalArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const DevMechLayout& Lay = A.lay;
  for (int i = threadIdx.x * 16; i < Lay.lds_bytes; i += blockDim.x * 16)
    *reinterpret_cast<uint4*>(smem + i) = *reinterpret_cast<const uint4*>(A.mech + i);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wpb = blockDim.x >> 6;
  Tabs t = make_tabs(smem, A.mech, Lay);
  const int wl_doubles = wave_lds_doubles(Lay.R, Lay.n_partials);
  WaveLds w = make_wave_lds(reinterpret_cast<double*>(smem + Lay.lds_bytes) + wave * wl_doubles, Lay.R, Lay.n_partials);
  const int64_t slot = static_cast<int64_t>(blockIdx.x) * wpb + wave;
  WaveScratch g = make_wave_scratch(A.scratch + slot * A.scratch_stride, Lay);
  if (lane < 16) g.coefbuf[lane] = t.coef[lane];
  for (int64_t c = slot; c < A.ncells; c += static_cast<int64_t>(gridDim.x) * wpb) {
    double y = lane < NS ? A.y[c * A.ld + lane] : 0.0;
    double P = A.P[c];
    if constexpr (MODE == 0) {
      double f = rhs(t, w, y, P, lane, NS);
      if (lane < NS) A.out[c * A.ld + lane] = f;
    } else {
      jacobian<NS>(t, w, g, y, P, lane);
      if (lane < NS)
        for (int j = 0; j < NS; ++j) A.out[(c * NS + lane) * NS + j] = g.J[j * NS + lane];
    }
  }
}

template <int NS, typename F>
__global__ __launch_bounds__(64) void lu_kernel(LuArgs A) {
  const int lane = threadIdx.x & 63;
  const int64_t c = blockIdx.x;
  LuReg<NS, F> L;
#pragma unroll
  for (int j = 0; j < NS; ++j) L.a[j] = lane < NS ? static_cast<F>(A.M[(c * NS + lane) * NS + j]) : F(0);
  bool ok;
  F x;
  F b = lane < NS ? static_cast<F>(A.b[c * NS + lane]) : F(0);
  if constexpr (sizeof(F) == 4) {
    __shared__ float buf[64];
    ok = gj_invert<NS>(L, lane);
    x = gj_apply<NS>(L, b, lane, buf);
  } else {
    ok = lu_factor<NS>(L, lane);
    x = lu_solve<NS>(L, b, lane);
  }
  if (lane < NS) A.x[c * NS + lane] = static_cast<double>(x);
  if (lane == 0) A.ok[c] = ok ? 1 : 0;
}

}  // namespace dev

// ---------------------------------------------------------------------------
// launch wrappers (host)
// ---------------------------------------------------------------------------
// Kernel variants: 0 FP64 LU; 1 FP32 explicit inverse (Gauss-Jordan); 2 FP32 in-place LU.
template <int NS, bool PROF>
static const void* bdf_fn_p(int variant) {
  if (variant == 1) return reinterpret_cast<const void*>(&dev::bdf_kernel<NS, true, PROF, true>);
  if (variant == 2) return reinterpret_cast<const void*>(&dev::bdf_kernel<NS, true, PROF, false>);
  return reinterpret_cast<const void*>(&dev::bdf_kernel<NS, false, PROF, false>);
}

template <int NS>
static const void* bdf_fn(int variant, int prof) {
  return prof ? bdf_fn_p<NS, true>(variant) : bdf_fn_p<NS, false>(variant);
}

template <int NS, bool PROF>
static void launch_bdf_p(const BdfArgs& A, int variant, int grid, int block, size_t shmem, hipStream_t st) {
  if (variant == 1)
    hipLaunchKernelGGL((dev::bdf_kernel<NS, true, PROF, true>), dim3(grid), dim3(block), shmem, st, A);
  else if (variant == 2)
    hipLaunchKernelGGL((dev::bdf_kernel<NS, true, PROF, false>), dim3(grid), dim3(block), shmem, st, A);
  else
    hipLaunchKernelGGL((dev::bdf_kernel<NS, false, PROF, false>), dim3(grid), dim3(block), shmem, st, A);
}

template <int NS>
static hipError_t launch_bdf_ns(const BdfArgs& A, int variant, int grid, int block, size_t shmem, hipStream_t st) {
  if (A.prof != nullptr)
    launch_bdf_p<NS, true>(A, variant, grid, block, shmem, st);
  else
    launch_bdf_p<NS, false>(A, variant, grid, block, shmem, st);
  return hipGetLastError();
}

#define HPO_FOR_EACH_NS(X) X(9) X(10) X(53)

int bdf_wave_lds_doubles(const DevMechLayout& L) { return dev::wave_lds_doubles(L.R, L.n_partials) + dev::kScalDoubles; }
int eval_wave_lds_doubles(const DevMechLayout& L) { return dev::wave_lds_doubles(L.R, L.n_partials); }
int64_t wave_scratch_doubles_host(const DevMechLayout& L) { return dev::wave_scratch_doubles(L); }

bool kernel_supports(int ns) {
#define HPO_CASE(N) \
  if (ns == N) return true;
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return false;
}

const void* bdf_kernel_fn(int ns, int variant, int prof) {
#define HPO_CASE(N) \
  if (ns == N) return bdf_fn<N>(variant, prof);
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return nullptr;
}

namespace dev {
// Zeroes a launch's work-queue counter on the launch's stream. A kernel rather than a memset, so
// that a caller capturing the launch into a HIP graph records two kernel nodes.
__global__ __launch_bounds__(64) void reset_counter(uint32_t* q) {
  if (threadIdx.x == 0) q[0] = 0u;
}
}  // namespace dev

hipError_t launch_reset_counter(uint32_t* q, hipStream_t st) {
  hipLaunchKernelGGL(dev::reset_counter, dim3(1), dim3(64), 0, st, q);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fixed device scratch of the no-allocation entry point (hpo_integrate_device_noalloc): the
// work-queue counter in its first 256 bytes, then the per-wave slots. One copy per GPU, held
// until the process ends. In host code `scratch` names a host-side stand-in, not device
// memory: only the address hipGetSymbolAddress returns for the current device is used.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];

namespace {
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoallocMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, void** ptr) {
  if (device < 0 || device >= kNoallocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_ptr[device] = p;
  }
  *ptr = g_noalloc_ptr[device];
  return hipSuccess;
}

hipError_t launch_bdf(int ns, const BdfArgs& A, int variant, int grid, int block, size_t shmem, hipStream_t st) {
#define HPO_CASE(N) \
  if (ns == N) return launch_bdf_ns<N>(A, variant, grid, block, shmem, st);
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_eval(int ns, const EvalArgs& A, int grid, int block, size_t shmem, hipStream_t st) {
#define HPO_CASE(N)                                                                         \
  if (ns == N) {                                                                            \
    if (A.mode == 0)                                                                        \
      hipLaunchKernelGGL((dev::eval_kernel<N, 0>), dim3(grid), dim3(block), shmem, st, A);  \
    else                                                                                    \
      hipLaunchKernelGGL((dev::eval_kernel<N, 1>), dim3(grid), dim3(block), shmem, st, A);  \
    return hipGetLastError();                                                               \
  }
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_lu(int ns, int fp32, const LuArgs& A, int64_t ncases, hipStream_t st) {
#define HPO_CASE(N)                                                                                        \
  if (ns == N) {                                                                                           \
    if (fp32)                                                                                              \
      hipLaunchKernelGGL((dev::lu_kernel<N, float>), dim3(static_cast<unsigned>(ncases)), dim3(64), 0, st, A);  \
    else                                                                                                   \
      hipLaunchKernelGGL((dev::lu_kernel<N, double>), dim3(static_cast<unsigned>(ncases)), dim3(64), 0, st, A); \
    return hipGetLastError();                                                                              \
  }
  HPO_FOR_EACH_NS(HPO_CASE)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

}  // namespace hpo
