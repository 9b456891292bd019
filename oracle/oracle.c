/*
 * CPU oracle (checker) for the stiff-chemistry BDF hot path. See oracle.h.
 *
 * TEST INFRASTRUCTURE ONLY. Parity status of each part:
 *   or_rhs, or_rates  pinned: pyJac golden vectors (GRI-3.0) + closed forms + invariants
 *   or_jac            pinned: pyJac golden Jacobian + central finite differences of or_rhs
 *   or_getrf/getrs    pinned: LAPACK dgetrf/dgetrs (scipy) on the same matrices
 *   or_cvode_*        pinned against closed forms (y'=lambda*y, linear systems), order of
 *                     convergence, and scipy BDF/Radau truth at rtol 1e-12; CVODE's exact
 *                     heuristic constants (cvHin, eta limits, Jacobian reuse) are restated
 *                     from the published SUNDIALS 6.x algorithm and are parity unpinned
 *                     against real SUNDIALS (its sources are not in the reference).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define R_GAS 8314.4621
#define P_ATM 101325.0

/* ------------------------------------------------------------------------- */
/* Thermodynamics: NASA7, per-species T_mid (chem_utils.c:194-514,1160-1480)  */
/* ------------------------------------------------------------------------- */
static void nasa7(const or_mech *m, int k, double T, double lnT, double *cp_R, double *h_RT,
                  double *s_R, double *dcp_R_dT) {
  const double *a = m->nasa + 14 * k + (T <= m->tmid[k] ? 0 : 7);
  *cp_R = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  *h_RT = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T;
  *s_R = a[0] * lnT + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6];
  if (dcp_R_dT) *dcp_R_dT = a[1] + T * (2 * a[2] + T * (3 * a[3] + T * 4 * a[4]));
}

static double karr(double lnA, double b, double th, double A, double lnT, double T) {
  if (b == 0.0 && th == 0.0) return A; /* pyJac emits the constant (rxn_rates.c:45) */
  return exp(lnA + b * lnT - th / T);
}

/* Pressure-dependent Arrhenius ("PLOG", Cantera's PlogRate): ln k is interpolated linearly in
 * ln P between the two levels that bracket the pressure, each level an Arrhenius expression;
 * below the first level or above the last, that level's rate. Also returns d ln k / dT at fixed
 * P. lv: n levels of (ln P, lnA, b, theta), ascending in P. */
static void plog_rate(const double *lv, int n, double lnP, double lnT, double T, double *k, double *dlnk) {
  int i = 0;
  double w = 0.0;
  if (lnP <= lv[0]) {
    i = 0;
    w = 0.0;
  } else if (lnP >= lv[4 * (n - 1)]) {
    i = n - 1;
    w = 0.0;
  } else {
    while (i + 1 < n && lv[4 * (i + 1)] <= lnP) ++i;
    w = (lnP - lv[4 * i]) / (lv[4 * (i + 1)] - lv[4 * i]);
  }
  const double *a = lv + 4 * i;
  double lk = a[1] + a[2] * lnT - a[3] / T;
  double dl = (a[2] + a[3] / T) / T;
  if (w != 0.0) {
    const double *c = a + 4;
    const double lk2 = c[1] + c[2] * lnT - c[3] / T;
    const double dl2 = (c[2] + c[3] / T) / T;
    lk += w * (lk2 - lk);
    dl += w * (dl2 - dl);
  }
  *k = exp(lk);
  *dlnk = dl;
}

static double ipow(double c, double nu) {
  if (nu == 1.0) return c;
  if (nu == 2.0) return c * c;
  if (nu == 3.0) return c * c * c;
  return pow(c, nu);
}

/* d(c^nu)/dc */
static double dipow(double c, double nu) {
  if (nu == 1.0) return 1.0;
  if (nu == 2.0) return 2.0 * c;
  if (nu == 3.0) return 3.0 * c * c;
  return nu * pow(c, nu - 1.0);
}

/* Everything the RHS and the Jacobian share. */
typedef struct {
  double T, lnT, P, rho, Wbar, m, yN, cpmix, dcpmix_dT;
  double *Y, *C, *cp_R, *h_RT, *g, *dcp_R_dT; /* [K] */
} state_t;

static void eval_state(const or_mech *m, double P, const double *y, state_t *s) {
  const int K = m->K;
  double T = y[0];
  s->T = T;
  s->lnT = log(T);
  s->P = P;
  double sumY = 0.0;
  for (int k = 0; k < K - 1; ++k) sumY += y[k + 1];
  s->yN = 1.0 - sumY; /* chem_utils.c:6-13 */
  double S = 0.0;
  for (int k = 0; k < K; ++k) {
    s->Y[k] = (k < K - 1) ? y[k + 1] : s->yN;
    S += s->Y[k] * (1.0 / m->W[k]);
  }
  s->Wbar = 1.0 / S;
  if (m->conv) {
    /* constant volume: the argument is the density (eval_conc_rho of dydt.c:143) */
    s->rho = P;
    s->m = P * S;
    s->P = s->m * R_GAS * T;
  } else {
    s->rho = P * s->Wbar / (R_GAS * T); /* chem_utils.c:42 */
    s->m = P / (R_GAS * T);              /* rxn_rates_pres_mod.c:20 */
  }
  s->cpmix = 0.0;
  s->dcpmix_dT = 0.0;
  for (int k = 0; k < K; ++k) {
    s->C[k] = s->rho * s->Y[k] * (1.0 / m->W[k]);
    double sR;
    nasa7(m, k, T, s->lnT, &s->cp_R[k], &s->h_RT[k], &sR, &s->dcp_R_dT[k]);
    s->g[k] = s->h_RT[k] - sR;
    s->cpmix += s->Y[k] * s->cp_R[k] * R_GAS / m->W[k];
    s->dcpmix_dT += s->Y[k] * s->dcp_R_dT[k] * R_GAS / m->W[k];
  }
}

/* Per-reaction rate pieces. */
typedef struct {
  double kf, kr, fwd, rev, pm;     /* rates and pressure modification */
  double dlnkf, dlnkr;             /* d ln k / dT at fixed C */
  double dpm_dM, dpm_dT;           /* d pm / d[M] (fixed T), d pm / dT (fixed [M]) */
  double M;                        /* third-body concentration */
} rxn_t;

static void eval_rxn(const or_mech *m, int r, const state_t *s, rxn_t *o) {
  const double T = s->T, lnT = s->lnT;
  const double *a = m->arr + 8 * r;
  const int nr = m->nreac[r], np = m->nprod[r];
  const int *rs = m->reac_sp + OR_MAX_SLOTS * r, *ps = m->prod_sp + OR_MAX_SLOTS * r;
  const double *rn = m->reac_nu + OR_MAX_SLOTS * r, *pn = m->prod_nu + OR_MAX_SLOTS * r;
  if (m->plog_ptr && m->plog_ptr[r + 1] > m->plog_ptr[r]) {
    plog_rate(m->plog + 4 * m->plog_ptr[r], m->plog_ptr[r + 1] - m->plog_ptr[r], log(s->P), lnT, T, &o->kf,
              &o->dlnkf);
  } else {
    o->kf = karr(a[0], a[1], a[2], a[3], lnT, T);
    o->dlnkf = (a[1] + a[2] / T) / T;
  }
  double fwd = o->kf;
  for (int i = 0; i < nr; ++i) fwd *= ipow(s->C[rs[i]], rn[i]);
  o->fwd = fwd;
  if (m->rev[r]) {
    /* Kc = (P_atm/(R T))^dnu exp(-sum nu g)  (rxn_rates.c:11-17) */
    double dG = 0.0, dnu = 0.0, dH = 0.0;
    for (int i = 0; i < np; ++i) {
      dG += pn[i] * s->g[ps[i]];
      dH += pn[i] * s->h_RT[ps[i]];
      dnu += pn[i];
    }
    for (int i = 0; i < nr; ++i) {
      dG -= rn[i] * s->g[rs[i]];
      dH -= rn[i] * s->h_RT[rs[i]];
      dnu -= rn[i];
    }
    double Kc = exp(dnu * log(P_ATM / (R_GAS * T)) - dG);
    o->kr = o->kf / Kc;
    /* d ln Kc/dT = (dH/RT - dnu)/T  (van 't Hoff, with the T^-dnu factor) */
    o->dlnkr = o->dlnkf - (dH - dnu) / T;
    double rev = o->kr;
    for (int i = 0; i < np; ++i) rev *= ipow(s->C[ps[i]], pn[i]);
    o->rev = rev;
  } else {
    o->kr = 0.0;
    o->dlnkr = 0.0;
    o->rev = 0.0;
  }
  o->pm = 1.0;
  o->dpm_dM = 0.0;
  o->dpm_dT = 0.0;
  o->M = 0.0;
  const int kind = m->kind[r];
  if (kind != 0) {
    double M = m->eff_base[r] * s->m;
    for (int e = m->eff_ptr[r]; e < m->eff_ptr[r + 1]; ++e) M += m->eff_val[e] * s->C[m->eff_sp[e]];
    o->M = M;
    if (kind == 1) {
      o->pm = M;
      o->dpm_dM = 1.0;
    } else {
      double k0 = karr(a[4], a[5], a[6], a[7], lnT, T);
      double dlnk0 = (a[5] + a[6] / T) / T;
      double kinf = o->kf;
      double Pr = k0 * M / kinf;
      double dPr_dT = Pr * (dlnk0 - o->dlnkf); /* at fixed [M] */
      double dPr_dM = k0 / kinf;
      double F = 1.0, dF_dPr = 0.0, dF_dT = 0.0;
      if (kind == 3) {
        const double *tr = m->troe + 5 * r;
        double ta = tr[0], T3 = tr[1], T1 = tr[2], T2 = tr[3];
        double e3 = exp(-T / T3), e1 = exp(-T / T1);
        double e2 = tr[4] != 0.0 ? exp(-T2 / T) : 0.0;
        double Fc = (1.0 - ta) * e3 + ta * e1 + e2;
        double dFc = -(1.0 - ta) / T3 * e3 - ta / T1 * e1 + (tr[4] != 0.0 ? T2 / (T * T) * e2 : 0.0);
        int fc_clamped = !(Fc > 1.0e-300);
        double logFc = log10(fmax(Fc, 1.0e-300)); /* rxn_rates_pres_mod.c:65 */
        int pr_clamped = !(Pr > 1.0e-300);
        double x = log10(fmax(Pr, 1.0e-300));
        double A = x - 0.67 * logFc - 0.4;
        double B = 0.806 - 1.1762 * logFc - 0.14 * x;
        double D = 1.0 + A * A / (B * B);
        double f = logFc / D;
        F = pow(10.0, f);
        /* df/dx and df/dlogFc */
        double dD_dA = 2.0 * A / (B * B), dD_dB = -2.0 * A * A / (B * B * B);
        double df_dx = -logFc / (D * D) * (dD_dA * 1.0 + dD_dB * (-0.14));
        double df_dc = 1.0 / D - logFc / (D * D) * (dD_dA * (-0.67) + dD_dB * (-1.1762));
        if (pr_clamped) df_dx = 0.0;
        double dc_dT = fc_clamped ? 0.0 : dFc / (Fc * log(10.0));
        /* F = 10^f: dF/dPr = F ln10 df/dx dx/dPr = F df/dx / Pr */
        dF_dPr = pr_clamped ? 0.0 : F * df_dx / Pr;
        dF_dT = F * log(10.0) * df_dc * dc_dT;
        /* combined below in a form without 1/Pr */
        double q = 1.0 / (1.0 + Pr);
        o->pm = F * Pr * q;
        double dpm_dPr = F * q * q + F * df_dx * q; /* = d/dPr (F Pr/(1+Pr)) */
        (void)dF_dPr;
        o->dpm_dM = dpm_dPr * dPr_dM;
        o->dpm_dT = dpm_dPr * dPr_dT + Pr * q * dF_dT;
      } else {
        double q = 1.0 / (1.0 + Pr);
        o->pm = Pr * q;
        double dpm_dPr = q * q;
        o->dpm_dM = dpm_dPr * dPr_dM;
        o->dpm_dT = dpm_dPr * dPr_dT;
        (void)F;
        (void)dF_dPr;
        (void)dF_dT;
      }
    }
  }
}

typedef struct {
  double *buf;
  state_t s;
} work_t;

static void work_init(work_t *w, int K) {
  w->buf = (double *)malloc(sizeof(double) * 6 * K);
  w->s.Y = w->buf;
  w->s.C = w->buf + K;
  w->s.cp_R = w->buf + 2 * K;
  w->s.h_RT = w->buf + 3 * K;
  w->s.g = w->buf + 4 * K;
  w->s.dcp_R_dT = w->buf + 5 * K;
}

static void work_free(work_t *w) { free(w->buf); }

void or_rates(const or_mech *m, double P, const double *y, double *conc, double *fwd,
              double *rev, double *presmod, double *wdot) {
  const int K = m->K;
  work_t w;
  work_init(&w, K);
  eval_state(m, P, y, &w.s);
  for (int k = 0; k < K; ++k) {
    conc[k] = w.s.C[k];
    wdot[k] = 0.0;
  }
  for (int r = 0; r < m->R; ++r) {
    rxn_t o;
    eval_rxn(m, r, &w.s, &o);
    fwd[r] = o.fwd;
    rev[r] = o.rev;
    presmod[r] = o.pm;
    double phi = (o.fwd - o.rev) * o.pm;
    for (int i = 0; i < m->nreac[r]; ++i)
      wdot[m->reac_sp[OR_MAX_SLOTS * r + i]] -= m->reac_nu[OR_MAX_SLOTS * r + i] * phi;
    for (int i = 0; i < m->nprod[r]; ++i)
      wdot[m->prod_sp[OR_MAX_SLOTS * r + i]] += m->prod_nu[OR_MAX_SLOTS * r + i] * phi;
  }
  work_free(&w);
}

/* dydt, constant pressure (dydt.c:7-133) */
void or_rhs(const or_mech *m, double P, const double *y, double *dy) {
  const int K = m->K;
  work_t w;
  work_init(&w, K);
  double *wdot = (double *)calloc(K, sizeof(double));
  eval_state(m, P, y, &w.s);
  for (int r = 0; r < m->R; ++r) {
    rxn_t o;
    eval_rxn(m, r, &w.s, &o);
    double phi = (o.fwd - o.rev) * o.pm;
    for (int i = 0; i < m->nreac[r]; ++i)
      wdot[m->reac_sp[OR_MAX_SLOTS * r + i]] -= m->reac_nu[OR_MAX_SLOTS * r + i] * phi;
    for (int i = 0; i < m->nprod[r]; ++i)
      wdot[m->prod_sp[OR_MAX_SLOTS * r + i]] += m->prod_nu[OR_MAX_SLOTS * r + i] * phi;
  }
  /* constant pressure: Q = sum_k H_k wdot_k (molar enthalpy H_k = h_RT R T) over rho cp;
     constant volume: the internal energies U_k = H_k - R T over rho cv, cv = cp - R / Wbar */
  const double cv = (double)m->conv;
  double Q = 0.0;
  for (int k = 0; k < K; ++k) Q += wdot[k] * (w.s.h_RT[k] - cv) * R_GAS * w.s.T;
  dy[0] = -Q / (w.s.rho * (w.s.cpmix - cv * R_GAS / w.s.Wbar));
  for (int k = 0; k < K - 1; ++k) dy[k + 1] = wdot[k] * m->W[k] / w.s.rho;
  free(wdot);
  work_free(&w);
}

/* Exact Jacobian of or_rhs, column-major (jacob.c:3-371 computes the same object). */
void or_jac(const or_mech *m, double P, const double *y, double *J) {
  const int K = m->K, n = K, R = m->R;
  work_t w;
  work_init(&w, K);
  eval_state(m, P, y, &w.s);
  const state_t *s = &w.s;
  double *JC = (double *)calloc((size_t)K * K, sizeof(double)); /* JC[k*K+i] = d wdot_k / d C_i */
  double *Jm = (double *)calloc(K, sizeof(double));             /* d wdot_k / d m */
  double *JT = (double *)calloc(K, sizeof(double));             /* d wdot_k / dT | C,m */
  double *wdot = (double *)calloc(K, sizeof(double));
  double *dphi = (double *)malloc(sizeof(double) * K);          /* d phi_r / d C_i (dense scratch) */
  for (int r = 0; r < R; ++r) {
    rxn_t o;
    eval_rxn(m, r, s, &o);
    const int nr = m->nreac[r], np = m->nprod[r];
    const int *rs = m->reac_sp + OR_MAX_SLOTS * r, *ps = m->prod_sp + OR_MAX_SLOTS * r;
    const double *rn = m->reac_nu + OR_MAX_SLOTS * r, *pn = m->prod_nu + OR_MAX_SLOTS * r;
    double dq = o.fwd - o.rev;
    double phi = dq * o.pm;
    for (int i = 0; i < K; ++i) dphi[i] = 0.0;
    for (int a = 0; a < nr; ++a) { /* d fwd / d C_rs[a] */
      double d = o.kf;
      for (int b = 0; b < nr; ++b) d *= (b == a) ? dipow(s->C[rs[b]], rn[b]) : ipow(s->C[rs[b]], rn[b]);
      dphi[rs[a]] += o.pm * d;
    }
    if (m->rev[r]) {
      for (int a = 0; a < np; ++a) {
        double d = o.kr;
        for (int b = 0; b < np; ++b) d *= (b == a) ? dipow(s->C[ps[b]], pn[b]) : ipow(s->C[ps[b]], pn[b]);
        dphi[ps[a]] -= o.pm * d;
      }
    }
    double dphi_dm = 0.0;
    if (m->kind[r] != 0) {
      double fM = dq * o.dpm_dM;
      for (int e = m->eff_ptr[r]; e < m->eff_ptr[r + 1]; ++e) dphi[m->eff_sp[e]] += fM * m->eff_val[e];
      dphi_dm = fM * m->eff_base[r];
    }
    double dphi_dT = o.pm * (o.fwd * o.dlnkf - o.rev * o.dlnkr) + dq * o.dpm_dT;
    /* scatter nu_kr * (...) */
    for (int side = 0; side < 2; ++side) {
      int cnt = side ? np : nr;
      const int *sp = side ? ps : rs;
      const double *nu = side ? pn : rn;
      double sg = side ? 1.0 : -1.0;
      for (int a = 0; a < cnt; ++a) {
        int k = sp[a];
        double c = sg * nu[a];
        wdot[k] += c * phi;
        Jm[k] += c * dphi_dm;
        JT[k] += c * dphi_dT;
        for (int i = 0; i < K; ++i) JC[(size_t)k * K + i] += c * dphi[i];
      }
    }
  }
  /* chain rule to y = [T, Y_0..Y_{K-2}] with Y_{K-1} = 1 - sum.
     Constant pressure: rho = P Wbar / (R T) and m = P / (R T) move with T and Y, so
     dC_i/dT = -C_i/T, dm/dT = -m/T, dC_i/dY_j carries -Wbar d_j C_i, and f_k = W_k wdot_k / rho
     picks up rho's derivatives. Constant volume: rho is fixed, C_i = rho Y_i / W_i does not
     depend on T, and m = rho sum_k Y_k / W_k gives dm/dY_j = rho d_j. d_j = 1/W_j - 1/W_{K-1}. */
  const int conv = m->conv;
  const double rho = s->rho, Wbar = s->Wbar, T = s->T;
  const double invWN = 1.0 / m->W[K - 1];
  double *sigma = (double *)calloc(K, sizeof(double));
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < K; ++i) sigma[k] += JC[(size_t)k * K + i] * s->C[i];
  /* dwdot_k/dY_j and dwdot_k/dT */
  double *dW = (double *)malloc(sizeof(double) * K * n); /* dW[k*n + col] */
  for (int k = 0; k < K; ++k) {
    dW[k * n + 0] = conv ? JT[k] : JT[k] - sigma[k] / T - (s->m / T) * Jm[k];
    for (int j = 0; j < K - 1; ++j) {
      double dj = 1.0 / m->W[j] - invWN;
      double d = JC[(size_t)k * K + j] * rho / m->W[j] - JC[(size_t)k * K + K - 1] * rho * invWN;
      dW[k * n + j + 1] = conv ? d + rho * dj * Jm[k] : d - Wbar * dj * sigma[k];
    }
  }
  /* species rows f_k = W_k wdot_k / rho */
  for (int k = 0; k < K - 1; ++k) {
    double fac = m->W[k] / rho;
    J[(k + 1) + 0 * n] = fac * (dW[k * n + 0] + (conv ? 0.0 : wdot[k] / T));
    for (int j = 0; j < K - 1; ++j) {
      double dj = 1.0 / m->W[j] - invWN;
      J[(k + 1) + (j + 1) * n] = fac * (dW[k * n + j + 1] + (conv ? 0.0 : wdot[k] * Wbar * dj));
    }
  }
  /* temperature row f_T = -Q/D: constant pressure Q = sum_k H_k wdot_k, D = rho cp; constant
     volume Q = sum_k U_k wdot_k with U_k = H_k - R T (dU_k/dT = cp_k - R), D = rho cv */
  const double cv = (double)conv;
  double Q = 0.0;
  for (int k = 0; k < K; ++k) Q += (s->h_RT[k] - cv) * R_GAS * T * wdot[k];
  const double cmix = s->cpmix - cv * R_GAS / Wbar;
  double D = rho * cmix;
  double fT = -Q / D;
  {
    double dQ = 0.0;
    for (int k = 0; k < K; ++k)
      dQ += (s->cp_R[k] - cv) * R_GAS * wdot[k] + (s->h_RT[k] - cv) * R_GAS * T * dW[k * n + 0];
    double dD = (conv ? 0.0 : -rho * s->cpmix / T) + rho * s->dcpmix_dT;
    J[0] = (-dQ - fT * dD) / D;
  }
  const double cN = (s->cp_R[K - 1] - cv) * R_GAS * invWN;
  for (int j = 0; j < K - 1; ++j) {
    double dj = 1.0 / m->W[j] - invWN;
    double dQ = 0.0;
    for (int k = 0; k < K; ++k) dQ += (s->h_RT[k] - cv) * R_GAS * T * dW[k * n + j + 1];
    double cj = (s->cp_R[j] - cv) * R_GAS / m->W[j];
    double dD = (conv ? 0.0 : -rho * Wbar * dj * s->cpmix) + rho * (cj - cN);
    J[0 + (j + 1) * n] = (-dQ - fT * dD) / D;
  }
  free(dW);
  free(sigma);
  free(dphi);
  free(wdot);
  free(JT);
  free(Jm);
  free(JC);
  work_free(&w);
}

/* ------------------------------------------------------------------------- */
/* Dense LU (SUNDIALS SUNDlsMat_denseGETRF / denseGETRS semantics)            */
/* ------------------------------------------------------------------------- */
long or_getrf(long n, double *a, long *p) {
  for (long k = 0; k < n; ++k) {
    double *colk = a + k * n;
    long l = k;
    for (long i = k + 1; i < n; ++i)
      if (fabs(colk[i]) > fabs(colk[l])) l = i;
    p[k] = l;
    if (colk[l] == 0.0) return k + 1;
    if (l != k)
      for (long j = 0; j < n; ++j) {
        double t = a[j * n + l];
        a[j * n + l] = a[j * n + k];
        a[j * n + k] = t;
      }
    double mult = 1.0 / colk[k];
    for (long i = k + 1; i < n; ++i) colk[i] *= mult;
    for (long j = k + 1; j < n; ++j) {
      double *colj = a + j * n;
      double akj = colj[k];
      if (akj != 0.0)
        for (long i = k + 1; i < n; ++i) colj[i] -= akj * colk[i];
    }
  }
  return 0;
}

void or_getrs(long n, const double *a, const long *p, double *b) {
  for (long k = 0; k < n; ++k) {
    long pk = p[k];
    if (pk != k) {
      double t = b[k];
      b[k] = b[pk];
      b[pk] = t;
    }
  }
  for (long k = 0; k < n - 1; ++k) {
    const double *colk = a + k * n;
    for (long i = k + 1; i < n; ++i) b[i] -= colk[i] * b[k];
  }
  for (long k = n - 1; k > 0; --k) {
    const double *colk = a + k * n;
    b[k] /= colk[k];
    for (long i = 0; i < k; ++i) b[i] -= colk[i] * b[k];
  }
  b[0] /= a[0];
}

static long getrf32(long n, float *a, long *p) {
  for (long k = 0; k < n; ++k) {
    float *colk = a + k * n;
    long l = k;
    for (long i = k + 1; i < n; ++i)
      if (fabsf(colk[i]) > fabsf(colk[l])) l = i;
    p[k] = l;
    if (colk[l] == 0.0f) return k + 1;
    if (l != k)
      for (long j = 0; j < n; ++j) {
        float t = a[j * n + l];
        a[j * n + l] = a[j * n + k];
        a[j * n + k] = t;
      }
    float mult = 1.0f / colk[k];
    for (long i = k + 1; i < n; ++i) colk[i] *= mult;
    for (long j = k + 1; j < n; ++j) {
      float *colj = a + j * n;
      float akj = colj[k];
      if (akj != 0.0f)
        for (long i = k + 1; i < n; ++i) colj[i] -= akj * colk[i];
    }
  }
  return 0;
}

static void getrs32(long n, const float *a, const long *p, float *b) {
  for (long k = 0; k < n; ++k) {
    long pk = p[k];
    if (pk != k) {
      float t = b[k];
      b[k] = b[pk];
      b[pk] = t;
    }
  }
  for (long k = 0; k < n - 1; ++k) {
    const float *colk = a + k * n;
    for (long i = k + 1; i < n; ++i) b[i] -= colk[i] * b[k];
  }
  for (long k = n - 1; k > 0; --k) {
    const float *colk = a + k * n;
    b[k] /= colk[k];
    for (long i = 0; i < k; ++i) b[i] -= colk[i] * b[k];
  }
  b[0] /= a[0];
}

/* ------------------------------------------------------------------------- */
/* CVODE BDF restatement (SUNDIALS 6.x cvode.c / cvode_ls.c / sunnonlinsol_newton.c) */
/* ------------------------------------------------------------------------- */
#define QMAX 5
#define L_MAX (QMAX + 1)
#define ADDON 1.0e-6
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define ETAMX1 10000.0
#define ETAMX2 10.0
#define ETAMX3 10.0
#define ETAMXF 0.2
#define ETAMIN 0.1
#define ETACF 0.25
#define THRESH 1.5
#define SMALL_NST 10
#define MXNCF 10
#define MXNEF 7
#define MXNEF1 3
#define SMALL_NEF 2
#define LONG_WAIT 10
#define NLS_MAXCOR 3
#define CRDOWN 0.3
#define DGMAX_LSETUP 0.3
#define RDIV 2.0
#define MSBP 20
#define MSBJ 51
#define DGMAX_JBAD 0.2
#define NLSCOEF 0.1
#define HLB_FACTOR 100.0
#define HUB_FACTOR 0.1
#define H_BIAS 0.5
#define MAX_ITERS 4
#define UROUND DBL_EPSILON

enum { FIRST_CALL = 0, PREV_CONV_FAIL = 1, PREV_ERR_FAIL = 2 };
enum { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };
enum { DO_ERROR_TEST = 2, PREDICT_AGAIN = 3, TRY_AGAIN = 5, CONV_RECVR = 901 };

typedef struct {
  int n;
  or_rhs_fn f;
  or_jac_fn jac;
  void *ud;
  or_opts o;
  double *zn[L_MAX + 1];
  double *ewt, *acor, *tempv, *ftemp, *y, *delta, *savedJ, *M, *res;
  float *M32, *b32;
  long *piv;
  double tn, h, hprime, hscale, eta, etamax, hu, next_h, etaq, etaqm1, etaqp1;
  double tau[L_MAX + 1], tq[6], l[L_MAX];
  double gamma, gammap, gamrat, rl1, crate, delp, acnrm, saved_tq5;
  int q, qprime, qwait, L, indx_acor, jcur, convfail;
  long nstlp, nstlj;
  or_stats st;
} cvmem;

static double wrms(int n, const double *v, const double *w) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    double t = v[i] * w[i];
    s += t * t;
  }
  return sqrt(s / n);
}

static int set_ewt(cvmem *c, const double *y, double *ewt) {
  for (int i = 0; i < c->n; ++i) {
    double t = c->o.rtol * fabs(y[i]) + c->o.atol;
    if (t <= 0.0) return -1;
    ewt[i] = 1.0 / t;
  }
  return 0;
}

static void rescale(cvmem *c) {
  double f = c->eta;
  for (int j = 1; j <= c->q; ++j) {
    for (int i = 0; i < c->n; ++i) c->zn[j][i] *= f;
    f *= c->eta;
  }
  c->h = c->hscale * c->eta;
  c->next_h = c->h;
  c->hscale = c->h;
}

static void predict(cvmem *c) {
  c->tn += c->h;
  for (int k = 1; k <= c->q; ++k)
    for (int j = c->q; j >= k; --j)
      for (int i = 0; i < c->n; ++i) c->zn[j - 1][i] += c->zn[j][i];
}

static void restore(cvmem *c, double saved_t) {
  c->tn = saved_t;
  for (int k = 1; k <= c->q; ++k)
    for (int j = c->q; j >= k; --j)
      for (int i = 0; i < c->n; ++i) c->zn[j - 1][i] -= c->zn[j][i];
}

static void increase_bdf(cvmem *c) {
  double alpha0, alpha1, prod, xi, xiold, hsum, A1;
  for (int i = 0; i <= QMAX; ++i) c->l[i] = 0.0;
  c->l[2] = alpha1 = prod = xiold = 1.0;
  alpha0 = -1.0;
  hsum = c->hscale;
  if (c->q > 1) {
    for (int j = 1; j < c->q; ++j) {
      hsum += c->tau[j + 1];
      xi = hsum / c->hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; --i) c->l[i] = c->l[i] * xiold + c->l[i - 1];
      xiold = xi;
    }
  }
  A1 = (-alpha0 - alpha1) / prod;
  for (int i = 0; i < c->n; ++i) c->zn[c->L][i] = A1 * c->zn[c->indx_acor][i];
  for (int j = 2; j <= c->q; ++j)
    for (int i = 0; i < c->n; ++i) c->zn[j][i] += c->l[j] * c->zn[c->L][i];
}

static void decrease_bdf(cvmem *c) {
  double hsum, xi;
  for (int i = 0; i <= QMAX; ++i) c->l[i] = 0.0;
  c->l[2] = 1.0;
  hsum = 0.0;
  for (int j = 1; j <= c->q - 2; ++j) {
    hsum += c->tau[j];
    xi = hsum / c->hscale;
    for (int i = j + 2; i >= 2; --i) c->l[i] = c->l[i] * xi + c->l[i - 1];
  }
  for (int j = 2; j < c->q; ++j)
    for (int i = 0; i < c->n; ++i) c->zn[j][i] = -c->l[j] * c->zn[c->q][i] + c->zn[j][i];
}

static void adjust_order(cvmem *c, int deltaq) {
  if (c->q == 2 && deltaq != 1) return;
  if (deltaq == 1) increase_bdf(c);
  else if (deltaq == -1) decrease_bdf(c);
}

static void adjust_params(cvmem *c) {
  if (c->qprime != c->q) {
    adjust_order(c, c->qprime - c->q);
    c->q = c->qprime;
    c->L = c->q + 1;
    c->qwait = c->L;
  }
  rescale(c);
}

static void set_tq_bdf(cvmem *c, double hsum, double alpha0, double alpha0_hat, double xi_inv,
                       double xistar_inv) {
  double A1, A2, A3, A4, A5, A6, C, Cpinv, Cppinv;
  A1 = 1.0 - alpha0_hat + alpha0;
  A2 = 1.0 + c->q * A1;
  c->tq[2] = fabs(A1 / (alpha0 * A2));
  c->tq[5] = fabs(A2 * xistar_inv / (c->l[c->q] * xi_inv));
  if (c->qwait == 1) {
    if (c->q > 1) {
      C = xistar_inv / c->l[c->q];
      A3 = alpha0 + 1.0 / c->q;
      A4 = alpha0_hat + xi_inv;
      Cpinv = (1.0 - A4 + A3) / A3;
      c->tq[1] = fabs(C * Cpinv);
    } else {
      c->tq[1] = 1.0;
    }
    hsum += c->tau[c->q];
    xi_inv = c->h / hsum;
    A5 = alpha0 - (1.0 / (c->q + 1));
    A6 = alpha0_hat - xi_inv;
    Cppinv = (1.0 - A6 + A5) / A2;
    c->tq[3] = fabs(Cppinv / (xi_inv * (c->q + 2) * A5));
  }
  c->tq[4] = NLSCOEF / c->tq[2];
}

static void set_bdf(cvmem *c) {
  double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
  c->l[0] = c->l[1] = xi_inv = xistar_inv = 1.0;
  for (int i = 2; i <= c->q; ++i) c->l[i] = 0.0;
  alpha0 = alpha0_hat = -1.0;
  hsum = c->h;
  if (c->q > 1) {
    for (int j = 2; j < c->q; ++j) {
      hsum += c->tau[j - 1];
      xi_inv = c->h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; --i) c->l[i] += c->l[i - 1] * xi_inv;
    }
    alpha0 -= 1.0 / c->q;
    xistar_inv = -c->l[1] - alpha0;
    hsum += c->tau[c->q - 1];
    xi_inv = c->h / hsum;
    alpha0_hat = -c->l[1] - xi_inv;
    for (int i = c->q; i >= 1; --i) c->l[i] += c->l[i - 1] * xistar_inv;
  }
  set_tq_bdf(c, hsum, alpha0, alpha0_hat, xi_inv, xistar_inv);
  c->rl1 = 1.0 / c->l[1];
  c->gamma = c->h * c->rl1;
  if (c->st.nst == 0) c->gammap = c->gamma;
  c->gamrat = (c->st.nst > 0) ? c->gamma / c->gammap : 1.0;
}

/* Newton residual: y = zn0 + ycor; res = rl1*zn1 + ycor - gamma*f(y) */
static int nls_residual(cvmem *c, const double *ycor, double *res) {
  for (int i = 0; i < c->n; ++i) c->y[i] = c->zn[0][i] + ycor[i];
  int rv = c->f(c->tn, c->y, c->ftemp, c->ud);
  c->st.nfe++;
  if (rv != 0) return -8; /* RHS failure: unrecoverable in this restatement */
  for (int i = 0; i < c->n; ++i) res[i] = c->rl1 * c->zn[1][i] + ycor[i] - c->gamma * c->ftemp[i];
  return 0;
}

/* cvLsSetup + cvLsLinSys + dense LU */
static int lsetup(cvmem *c) {
  const int n = c->n;
  double dgamma = fabs(c->gamma / c->gammap - 1.0);
  int jbad = (c->st.nst == 0) || (c->st.nst >= c->nstlj + MSBJ) ||
             (c->convfail == FAIL_BAD_J && dgamma < DGMAX_JBAD) || (c->convfail == FAIL_OTHER);
  if (!jbad) {
    c->jcur = 0;
    memcpy(c->M, c->savedJ, sizeof(double) * n * n);
  } else {
    c->jcur = 1;
    memset(c->M, 0, sizeof(double) * n * n);
    int rv = c->jac(c->tn, c->y, c->ftemp, c->M, c->ud);
    c->st.nje++;
    c->nstlj = c->st.nst;
    if (rv != 0) return -1;
    memcpy(c->savedJ, c->M, sizeof(double) * n * n);
  }
  /* M = I - gamma J */
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) c->M[j * n + i] = -c->gamma * c->M[j * n + i] + (i == j ? 1.0 : 0.0);
  c->st.nlu++;
  long ier;
  if (c->o.lu_fp32) {
    for (int k = 0; k < n * n; ++k) c->M32[k] = (float)c->M[k];
    ier = getrf32(n, c->M32, c->piv);
    /* keep the FP64 M for the residual of the refinement sweeps */
  } else {
    ier = or_getrf(n, c->M, c->piv);
  }
  /* cvNlsLSetup bookkeeping */
  c->st.nsetups++;
  c->gamrat = 1.0;
  c->gammap = c->gamma;
  c->crate = 1.0;
  c->nstlp = c->st.nst;
  if (ier > 0) return 1;
  return 0;
}

static void lsolve(cvmem *c, double *b) {
  const int n = c->n;
  c->st.nsolve++;
  if (!c->o.lu_fp32) {
    or_getrs(n, c->M, c->piv, b);
  } else {
    /* x0 = LU32 \ b ; for s in sweeps: r = b - M x ; x += LU32 \ r */
    double *x = c->res; /* scratch (res is not live during the solve) */
    for (int i = 0; i < n; ++i) c->b32[i] = (float)b[i];
    getrs32(n, c->M32, c->piv, c->b32);
    for (int i = 0; i < n; ++i) x[i] = c->b32[i];
    for (int s = 0; s < c->o.refine_sweeps; ++s) {
      for (int i = 0; i < n; ++i) {
        double r = b[i];
        for (int j = 0; j < n; ++j) r -= c->M[j * n + i] * x[j];
        c->b32[i] = (float)r;
      }
      getrs32(n, c->M32, c->piv, c->b32);
      for (int i = 0; i < n; ++i) x[i] += c->b32[i];
    }
    for (int i = 0; i < n; ++i) b[i] = x[i];
  }
  if (c->gamrat != 1.0)
    for (int i = 0; i < n; ++i) b[i] *= 2.0 / (1.0 + c->gamrat);
}

/* SUNNonlinSolSolve_Newton + cvNlsConvTest; returns 0, CONV_RECVR, or <0 */
static int nls_newton(cvmem *c, int callSetup) {
  const int n = c->n;
  double *ycor = c->acor, *delta = c->delta;
  int jbad = 0, retval;
  for (int i = 0; i < n; ++i) ycor[i] = 0.0;
  for (;;) {
    retval = nls_residual(c, ycor, delta);
    if (retval != 0) break;
    if (callSetup) {
      if (jbad) c->convfail = FAIL_BAD_J;
      retval = lsetup(c);
      if (retval < 0) break;
      if (retval > 0) {
        retval = CONV_RECVR;
        break;
      }
    }
    int m = 0;
    for (;;) {
      c->st.nni++;
      for (int i = 0; i < n; ++i) delta[i] = -delta[i];
      lsolve(c, delta);
      for (int i = 0; i < n; ++i) ycor[i] += delta[i];
      /* convergence test */
      double del = wrms(n, delta, c->ewt);
      if (m > 0) c->crate = fmax(CRDOWN * c->crate, del / c->delp);
      double dcon = del * fmin(1.0, c->crate) / c->tq[4];
      if (dcon <= 1.0) {
        c->acnrm = (m == 0) ? del : wrms(n, ycor, c->ewt);
        c->jcur = 0;
        return 0;
      }
      if (m >= 1 && del > RDIV * c->delp) {
        retval = CONV_RECVR;
        break;
      }
      c->delp = del;
      m++;
      if (m >= NLS_MAXCOR) {
        retval = CONV_RECVR;
        break;
      }
      retval = nls_residual(c, ycor, delta);
      if (retval != 0) break;
    }
    if (retval == CONV_RECVR && !c->jcur) {
      callSetup = 1;
      jbad = 1;
      for (int i = 0; i < n; ++i) ycor[i] = 0.0;
      continue;
    }
    break;
  }
  return retval;
}

/* cvNls: returns 0 on success, CONV_RECVR, or <0 */
static int cv_nls(cvmem *c, int nflag) {
  c->convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
  int callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (c->st.nst == 0) ||
                  (c->st.nst >= c->nstlp + MSBP) || (fabs(c->gamrat - 1.0) > DGMAX_LSETUP);
  int flag = nls_newton(c, callSetup);
  if (flag != 0) return flag;
  for (int i = 0; i < c->n; ++i) c->y[i] = c->zn[0][i] + c->acor[i];
  c->jcur = 0;
  return 0;
}

static int handle_nflag(cvmem *c, int *nflag, double saved_t, int *ncf) {
  if (*nflag == 0) return DO_ERROR_TEST;
  c->st.ncfn++;
  restore(c, saved_t);
  if (*nflag < 0) return OR_CONV_FAILURE;
  (*ncf)++;
  c->etamax = 1.0;
  if (*ncf == MXNCF) return OR_CONV_FAILURE;
  c->eta = ETACF;
  *nflag = PREV_CONV_FAIL;
  rescale(c);
  return PREDICT_AGAIN;
}

static int do_error_test(cvmem *c, int *nflag, double saved_t, int *nef, double *dsmp) {
  double dsm = c->acnrm * c->tq[2];
  *dsmp = dsm;
  if (dsm <= 1.0) return 0;
  (*nef)++;
  c->st.netf++;
  *nflag = PREV_ERR_FAIL;
  restore(c, saved_t);
  if (*nef == MXNEF) return OR_ERR_FAILURE;
  c->etamax = 1.0;
  if (*nef <= MXNEF1) {
    c->eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / c->L) + ADDON);
    c->eta = fmax(ETAMIN, c->eta);
    if (*nef >= SMALL_NEF) c->eta = fmin(c->eta, ETAMXF);
    rescale(c);
    return TRY_AGAIN;
  }
  if (c->q > 1) {
    c->eta = ETAMIN;
    adjust_order(c, -1);
    c->L = c->q;
    c->q--;
    c->qwait = c->L;
    rescale(c);
    return TRY_AGAIN;
  }
  c->eta = ETAMIN;
  c->h *= c->eta;
  c->next_h = c->h;
  c->hscale = c->h;
  c->qwait = LONG_WAIT;
  int rv = c->f(c->tn, c->zn[0], c->tempv, c->ud);
  c->st.nfe++;
  if (rv != 0) return OR_CONV_FAILURE;
  for (int i = 0; i < c->n; ++i) c->zn[1][i] = c->h * c->tempv[i];
  return TRY_AGAIN;
}

static void complete_step(cvmem *c) {
  c->st.nst++;
  c->hu = c->h;
  for (int i = c->q; i >= 2; --i) c->tau[i] = c->tau[i - 1];
  if (c->q == 1 && c->st.nst > 1) c->tau[2] = c->tau[1];
  c->tau[1] = c->h;
  for (int j = 0; j <= c->q; ++j)
    for (int i = 0; i < c->n; ++i) c->zn[j][i] += c->l[j] * c->acor[i];
  c->qwait--;
  if (c->qwait == 1 && c->q != QMAX) {
    for (int i = 0; i < c->n; ++i) c->zn[QMAX][i] = c->acor[i];
    c->saved_tq5 = c->tq[5];
    c->indx_acor = QMAX;
  }
}

static void set_eta(cvmem *c) {
  if (c->eta < THRESH) {
    c->eta = 1.0;
    c->hprime = c->h;
  } else {
    c->eta = fmin(c->eta, c->etamax);
    c->hprime = c->h * c->eta;
  }
}

static double compute_etaqm1(cvmem *c) {
  c->etaqm1 = 0.0;
  if (c->q > 1) {
    double ddn = wrms(c->n, c->zn[c->q], c->ewt) * c->tq[1];
    c->etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / c->q) + ADDON);
  }
  return c->etaqm1;
}

static double compute_etaqp1(cvmem *c) {
  c->etaqp1 = 0.0;
  if (c->q != QMAX) {
    if (c->saved_tq5 == 0.0) return c->etaqp1;
    double r = c->h / c->tau[2], pw = 1.0;
    for (int i = 0; i < c->L; ++i) pw *= r;
    double cquot = (c->tq[5] / c->saved_tq5) * pw;
    for (int i = 0; i < c->n; ++i) c->tempv[i] = -cquot * c->zn[QMAX][i] + c->acor[i];
    double dup = wrms(c->n, c->tempv, c->ewt) * c->tq[3];
    c->etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (c->L + 1)) + ADDON);
  }
  return c->etaqp1;
}

static void choose_eta(cvmem *c) {
  double etam = fmax(c->etaqm1, fmax(c->etaq, c->etaqp1));
  if (etam < THRESH) {
    c->eta = 1.0;
    c->qprime = c->q;
    return;
  }
  if (etam == c->etaqm1) {
    c->eta = c->etaqm1;
    c->qprime = c->q - 1;
  } else if (etam == c->etaq) {
    c->eta = c->etaq;
    c->qprime = c->q;
  } else {
    c->eta = c->etaqp1;
    c->qprime = c->q + 1;
    for (int i = 0; i < c->n; ++i) c->zn[QMAX][i] = c->acor[i];
  }
}

static void prepare_next_step(cvmem *c, double dsm) {
  if (c->etamax == 1.0) {
    c->qwait = c->qwait > 2 ? c->qwait : 2;
    c->qprime = c->q;
    c->hprime = c->h;
    c->eta = 1.0;
    return;
  }
  c->etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / c->L) + ADDON);
  if (c->qwait != 0) {
    c->eta = c->etaq;
    c->qprime = c->q;
    set_eta(c);
    return;
  }
  c->qwait = 2;
  compute_etaqm1(c);
  compute_etaqp1(c);
  choose_eta(c);
  set_eta(c);
}

static int cv_step(cvmem *c) {
  double saved_t = c->tn, dsm = 0.0;
  int ncf = 0, nef = 0, nflag = FIRST_CALL, kflag, eflag;
  if (c->st.nst > 0 && c->hprime != c->h) adjust_params(c);
  for (;;) {
    predict(c);
    set_bdf(c);
    nflag = cv_nls(c, nflag);
    kflag = handle_nflag(c, &nflag, saved_t, &ncf);
    if (kflag == PREDICT_AGAIN) continue;
    if (kflag != DO_ERROR_TEST) return kflag;
    eflag = do_error_test(c, &nflag, saved_t, &nef, &dsm);
    if (eflag == TRY_AGAIN) continue;
    if (eflag != 0) return eflag;
    break;
  }
  complete_step(c);
  prepare_next_step(c, dsm);
  c->etamax = (c->st.nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  for (int i = 0; i < c->n; ++i) c->acor[i] *= c->tq[2];
  return 0;
}

static double upper_bound_h0(cvmem *c, double tdist) {
  double hub_inv = 0.0;
  for (int i = 0; i < c->n; ++i) {
    double t1 = c->o.rtol * fabs(c->zn[0][i]) + c->o.atol; /* 1/ewt */
    t1 = HUB_FACTOR * fabs(c->zn[0][i]) + t1;
    double r = fabs(c->zn[1][i]) / t1;
    if (r > hub_inv) hub_inv = r;
  }
  double hub = HUB_FACTOR * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  return hub;
}

static int ydd_norm(cvmem *c, double hg, double *yddnrm) {
  for (int i = 0; i < c->n; ++i) c->y[i] = hg * c->zn[1][i] + c->zn[0][i];
  int rv = c->f(c->tn + hg, c->y, c->tempv, c->ud);
  c->st.nfe++;
  if (rv != 0) return -1;
  for (int i = 0; i < c->n; ++i) c->tempv[i] = (c->tempv[i] - c->zn[1][i]) * (1.0 / hg);
  *yddnrm = wrms(c->n, c->tempv, c->ewt);
  return 0;
}

static int cv_hin(cvmem *c, double tout) {
  double tdiff = tout - c->tn;
  if (tdiff == 0.0) return OR_TOO_CLOSE;
  int sign = tdiff > 0 ? 1 : -1;
  double tdist = fabs(tdiff);
  double tround = UROUND * fmax(fabs(c->tn), fabs(tout));
  if (tdist < 2.0 * tround) return OR_TOO_CLOSE;
  double hlb = HLB_FACTOR * tround;
  double hub = upper_bound_h0(c, tdist);
  double hg = sqrt(hlb * hub);
  if (hub < hlb) {
    c->h = sign == -1 ? -hg : hg;
    return 0;
  }
  double hnew = hg, yddnrm = 0.0;
  for (int count1 = 1; count1 <= MAX_ITERS; ++count1) {
    double hgs = hg * sign;
    if (ydd_norm(c, hgs, &yddnrm) != 0) return OR_CONV_FAILURE;
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    if (count1 == MAX_ITERS) break;
    double hrat = hnew / hg;
    if (hrat > 0.5 && hrat < 2.0) break;
    if (count1 > 1 && hrat > 2.0) {
      hnew = hg;
      break;
    }
    hg = hnew;
  }
  double h0 = H_BIAS * hnew;
  if (h0 < hlb) h0 = hlb;
  if (h0 > hub) h0 = hub;
  if (sign == -1) h0 = -h0;
  c->h = h0;
  return 0;
}

static void get_dky(cvmem *c, double t, double *dky) {
  double s = (t - c->tn) / c->h;
  /* N_VLinearCombination over zn[q], ..., zn[0] with c_j = s^j */
  double cj = 1.0;
  for (int i = 0; i < c->q; ++i) cj *= s;
  for (int i = 0; i < c->n; ++i) dky[i] = cj * c->zn[c->q][i];
  for (int j = c->q - 1; j >= 0; --j) {
    cj = 1.0;
    for (int i = 0; i < j; ++i) cj *= s;
    for (int i = 0; i < c->n; ++i) dky[i] += cj * c->zn[j][i];
  }
}

int or_cvode_generic(int n, or_rhs_fn f, or_jac_fn jac, void *ud, double t0, double tout,
                     double *yio, const or_opts *o, or_stats *st) {
  cvmem c;
  memset(&c, 0, sizeof(c));
  c.n = n;
  c.f = f;
  c.jac = jac;
  c.ud = ud;
  c.o = *o;
  size_t nv = (size_t)(L_MAX + 1) + 9;
  double *pool = (double *)calloc(nv * n + 2 * (size_t)n * n, sizeof(double));
  for (int j = 0; j <= L_MAX; ++j) c.zn[j] = pool + (size_t)j * n;
  double *p = pool + (size_t)(L_MAX + 1) * n;
  c.ewt = p; p += n;
  c.acor = p; p += n;
  c.tempv = p; p += n;
  c.ftemp = p; p += n;
  c.y = p; p += n;
  c.delta = p; p += n;
  c.res = p; p += n;
  p += 2 * n;
  c.savedJ = p; p += (size_t)n * n;
  c.M = p;
  c.M32 = (float *)calloc((size_t)n * n + n, sizeof(float));
  c.b32 = c.M32 + (size_t)n * n;
  c.piv = (long *)calloc(n, sizeof(long));

  int istate = OR_SUCCESS;
  /* CVodeInit */
  for (int i = 0; i < n; ++i) c.zn[0][i] = yio[i];
  c.tn = t0;
  c.q = 1;
  c.L = 2;
  c.qwait = c.L;
  c.etamax = ETAMX1;
  c.indx_acor = QMAX;
  c.crate = 1.0;
  if (set_ewt(&c, c.zn[0], c.ewt) != 0) {
    istate = OR_ILL_INPUT;
    goto done;
  }
  if (c.f(c.tn, c.zn[0], c.zn[1], c.ud) != 0) {
    istate = OR_CONV_FAILURE;
    goto done;
  }
  c.st.nfe++;
  {
    int hf = cv_hin(&c, tout);
    if (hf != 0) {
      istate = hf;
      goto done;
    }
  }
  c.hscale = c.h;
  c.hprime = c.h;
  for (int i = 0; i < n; ++i) c.zn[1][i] *= c.h;

  for (long nstloc = 0;; ) {
    c.next_h = c.h;
    if (c.st.nst > 0 && set_ewt(&c, c.zn[0], c.ewt) != 0) {
      istate = OR_ILL_INPUT;
      for (int i = 0; i < n; ++i) yio[i] = c.zn[0][i];
      break;
    }
    if (c.o.maxsteps > 0 && nstloc >= c.o.maxsteps) {
      istate = OR_TOO_MUCH_WORK;
      for (int i = 0; i < n; ++i) yio[i] = c.zn[0][i];
      break;
    }
    if (UROUND * wrms(n, c.zn[0], c.ewt) > 1.0) {
      istate = OR_TOO_MUCH_ACC;
      for (int i = 0; i < n; ++i) yio[i] = c.zn[0][i];
      break;
    }
    int kflag = cv_step(&c);
    if (kflag != 0) {
      istate = kflag;
      for (int i = 0; i < n; ++i) yio[i] = c.zn[0][i];
      break;
    }
    nstloc++;
    if ((c.tn - tout) * c.h >= 0.0) {
      get_dky(&c, tout, yio);
      break;
    }
  }
done:
  c.st.qlast = c.q;
  c.st.hlast = c.hu;
  if (st) *st = c.st;
  free(c.piv);
  free(c.M32);
  free(pool);
  return istate;
}

/* Chemistry callbacks */
typedef struct {
  const or_mech *m;
  double P;
} chem_ud;

static int chem_f(double t, const double *y, double *dy, void *ud) {
  (void)t;
  chem_ud *u = (chem_ud *)ud;
  or_rhs(u->m, u->P, y, dy);
  return 0;
}

static int chem_j(double t, const double *y, const double *fy, double *J, void *ud) {
  (void)t;
  (void)fy;
  chem_ud *u = (chem_ud *)ud;
  or_jac(u->m, u->P, y, J);
  return 0;
}

int or_cvode_cell(const or_mech *m, double P, double dt, double *y, const or_opts *o,
                  or_stats *st) {
  chem_ud u = {m, P};
  return or_cvode_generic(m->K, chem_f, chem_j, &u, 0.0, dt, y, o, st);
}

void or_omp_info(int *max_threads, int *sched_kind, int *chunk) {
#ifdef _OPENMP
  omp_sched_t k;
  int c;
  omp_get_schedule(&k, &c);
  *max_threads = omp_get_max_threads();
  *sched_kind = (int)k;
  *chunk = c;
#else
  *max_threads = 1;
  *sched_kind = 0;
  *chunk = 0;
#endif
}

void or_cvode_batch(const or_mech *m, long ncells, const double *P, const double *dt,
                    double *y, const or_opts *o, int *status, or_stats *stats, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const int K = m->K;
#pragma omp parallel for schedule(runtime)
  for (long c = 0; c < ncells; ++c) {
    or_stats st;
    int rc = or_cvode_cell(m, P[c], dt[c], y + (size_t)c * K, o, &st);
    if (status) status[c] = rc;
    if (stats) stats[c] = st;
  }
}
