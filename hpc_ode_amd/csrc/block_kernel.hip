// Block-per-cell persistent BDF kernel for gfx950: one 256-thread workgroup (four waves)
// integrates one cell, thread t owning state component t, so mechanisms with more species
// than a wave has lanes (nDodecane_Reitz, 100 species) run, and for smaller ones the
// reactions and the Newton matrix spread over four waves instead of one.
//
// The controller is the one of the wave kernel (bdf_cell.hpp); this file provides the
// workgroup versions of its operations:
//  * reductions: a wave sum (DPP) per wave, then the four wave partials through LDS, added in
//    a fixed order so every thread holds the same value;
//  * RHS: species thermo on threads 0..K-1, reactions strided over the 256 threads, species
//    production rates gathered per species from the compressed row list (fixed order);
//  * Jacobian: per reaction the derivative slots, per species row the owner-computed
//    d(wdot)/dC row, per column the enthalpy-weighted sum, then the chain rule to (T, Y) with
//    one Jacobian row per thread in the workgroup's scratch;
//  * Newton matrix: Gauss-Jordan inverse with LAPACK-style partial pivoting (largest candidate,
//    lowest row on ties), one register row per thread, the pivot row broadcast through LDS;
//    a solve is a matrix-vector product. FP64, or FP32 with FP64 residual refinement sweeps.
#include <hip/hip_runtime.h>

#include "bdf_cell.hpp"

#ifndef HPO_BLK_WAVES_PER_EU_FP64
#define HPO_BLK_WAVES_PER_EU_FP64 1
#endif
// Two waves per SIMD for the FP32 Newton matrix (a 100-float row fits the 256 registers):
// two workgroups per CU instead of one, 65,536 GRI cells in 2,452 ms against 3,995 ms and
// 16,384 nDodecane cells in 2,434 ms against 2,718 ms (first substep).
#ifndef HPO_BLK_WAVES_PER_EU_FP32
#define HPO_BLK_WAVES_PER_EU_FP32 2
#endif

namespace hpo {
namespace dev {

constexpr int kBlkWaves = 4;
constexpr int kBlkThreads = 64 * kBlkWaves;

struct BlkTabs {
  const double* sp;        // [K][kSpStride]
  const BlkRxn* rxn;       // [R]
  double* arr;             // [R][3] (in LDS: the PLOG rows are rewritten for each cell)
  const PdepRec* pd;
  const double* effv;
  const uint8_t* effs;
  const uint16_t* csr_ptr; // [K+1]
  const double* coef;      // distinct net stoichiometric coefficients
  const uint32_t* csr;     // r | coefficient index << 16
  const double* exptab;
  const uint8_t* slot_sp;  // global
  const uint32_t* col_ptr; // global
  const uint32_t* col;     // global
  const uint16_t* plog_rxn;   // global: device reaction of each PLOG rate
  const uint16_t* plog_ptr;   // global: its levels
  const double* plog_lv;      // global: [level][4] ln P, lnA, b, theta
  const uint16_t* lump_ptr;   // global: products of lumped reactions
  const uint32_t* lump;       // global: species | coefficient index << 16
  int K, R, n_plog;
  int conv;                // reactor: 0 constant pressure (P is the pressure), 1 constant volume
                           // (P is the density)
};

// Tables at offsets below lds_bytes were staged into LDS; the rest are read from global memory.
HPO_DI BlkTabs make_blk_tabs(uint8_t* lds, const uint8_t* glob, const BlkLayout& L) {
  auto at = [&](int32_t off) -> const uint8_t* { return off < L.lds_bytes ? lds + off : glob + off; };
  BlkTabs t;
  t.sp = reinterpret_cast<const double*>(at(L.off_species));
  t.rxn = reinterpret_cast<const BlkRxn*>(at(L.off_rxn));
  t.arr = reinterpret_cast<double*>(lds + L.off_arr);   // always staged
  t.pd = reinterpret_cast<const PdepRec*>(at(L.off_pdep));
  t.effv = reinterpret_cast<const double*>(at(L.off_eff_val));
  t.effs = at(L.off_eff_sp);
  t.csr_ptr = reinterpret_cast<const uint16_t*>(at(L.off_csr_ptr));
  t.coef = reinterpret_cast<const double*>(at(L.off_coef));
  t.csr = reinterpret_cast<const uint32_t*>(at(L.off_csr));
  t.exptab = reinterpret_cast<const double*>(lds + L.off_exptab);
  t.slot_sp = glob + L.off_slot_sp;
  t.col_ptr = reinterpret_cast<const uint32_t*>(glob + L.off_col_ptr);
  t.col = reinterpret_cast<const uint32_t*>(glob + L.off_col);
  t.plog_rxn = reinterpret_cast<const uint16_t*>(glob + L.off_plog_rxn);
  t.plog_ptr = reinterpret_cast<const uint16_t*>(glob + L.off_plog_ptr);
  t.plog_lv = reinterpret_cast<const double*>(glob + L.off_plog_lv);
  t.lump_ptr = reinterpret_cast<const uint16_t*>(glob + L.off_lump_ptr);
  t.lump = reinterpret_cast<const uint32_t*>(glob + L.off_lump);
  t.K = L.K;
  t.R = L.R;
  t.n_plog = L.n_plog;
  t.conv = 0;
  return t;
}

// PLOG rates at the cell's pressure P (Pa): each is an Arrhenius expression whose lnA, b and
// theta are those of the pressure level at or below P and the one above it, interpolated
// linearly in ln P (the interpolation of ln k is then exact for every T); below the first level
// or above the last, that level's. Written into the staged Arrhenius table; the caller orders
// the writes before the rates are read.
HPO_DI void blk_plog_cell(const BlkTabs& t, double P, int tid) {
  if (t.n_plog == 0) return;
  const double lnP = log(P);
  for (int i = tid; i < t.n_plog; i += kBlkThreads) {
    const int b = t.plog_ptr[i], e = t.plog_ptr[i + 1];
    const double* lv = t.plog_lv;
    int j = b;
    double w = 0.0;
    if (lnP > lv[4 * b] && lnP < lv[4 * (e - 1)]) {
      while (j + 2 < e && lv[4 * (j + 1)] <= lnP) ++j;
      w = (lnP - lv[4 * j]) / (lv[4 * (j + 1)] - lv[4 * j]);
    } else if (lnP >= lv[4 * (e - 1)]) {
      j = e - 1;
    }
    const double* a = lv + 4 * j;
    double lnA = a[1], bb = a[2], th = a[3];
    if (w != 0.0) {
      lnA += w * (a[5] - a[1]);
      bb += w * (a[6] - a[2]);
      th += w * (a[7] - a[3]);
    }
    double* dst = t.arr + 3 * t.plog_rxn[i];
    dst[0] = lnA;
    dst[1] = bb;
    dst[2] = th;
  }
}

// Workgroup LDS work area (after the tables). Sizes: NB = kBlkThreads.
struct BlkLds {
  double* y;      // [NB]   state broadcast
  double* vec;    // [NB]   vector broadcast (solves, RHS result)
  double* vec2;   // [NB]
  double* C;      // [K+1]  concentrations (index K: the dummy species, C = 1)
  double* g;      // [K+1]  G/RT (dummy: 0)
  double* hRT;    // [K+1]  H/RT (dummy: 0)
  double* cp;     // [K+1]  mass cp
  double* phi;    // [R]    net rates
  double* sp4;    // [5][K] Jacobian per-species values: wdot, JT, Jm, sigma, dQ/dC
  double* red;    // [8 * 8] reduction partials
  double* prow;   // pivot row [NB], or the pivot rows and multipliers of a block of elimination
                  // steps (blk_prow_doubles)
  int* perm;      // [NB]   pivot row of each Gauss-Jordan step
  int* iscal;     // [16]   pivot keys per wave, work-queue cell
};
// States of more than kBlkRegMax variables keep the Newton matrix in the workgroup's scratch
// (column j of the thread rows at M[j * kBlkThreads + t], so a column is one coalesced access)
// instead of one register row per thread.
constexpr int kBlkRegMax = 128;
// The pivot-row area, sized for FP64: with the matrix in registers, each wave's candidate
// pivot row for two steps (blk_gj_invert: 2 x 4 rows); with the matrix in scratch, the pivot
// rows of a block of kBlkGjSteps elimination steps and the multipliers of the current pivot
// row (blk_gj_invert_g).
constexpr int kBlkGjSteps = 8;
__host__ __device__ inline int blk_prow_doubles(int K) {
  const int rows = K > kBlkRegMax ? kBlkGjSteps * (K + 1) : 2 * 4 * K;
  return rows > kBlkThreads ? rows : kBlkThreads;
}
__host__ __device__ inline int blk_lds_doubles(int K, int R) {
  return 3 * kBlkThreads + 4 * (K + 2) + ((R + 1) & ~1) + 5 * ((K + 1) & ~1) + 64 + blk_prow_doubles(K) +
         kBlkThreads / 2 + 8;
}
HPO_DI BlkLds make_blk_lds(double* base, int K, int R) {
  BlkLds w;
  double* p = base;
  w.y = p; p += kBlkThreads;
  w.vec = p; p += kBlkThreads;
  w.vec2 = p; p += kBlkThreads;
  w.C = p; p += K + 2;
  w.g = p; p += K + 2;
  w.hRT = p; p += K + 2;
  w.cp = p; p += K + 2;
  w.phi = p; p += (R + 1) & ~1;
  w.sp4 = p; p += 5 * ((K + 1) & ~1);
  w.red = p; p += 64;
  w.prow = p; p += blk_prow_doubles(K);
  w.perm = reinterpret_cast<int*>(p); p += kBlkThreads / 2;
  w.iscal = reinterpret_cast<int*>(p);
  return w;
}

// Per-workgroup scratch (doubles): d(phi)/dC slots, per-reaction values, d(wdot)/dC rows,
// the Jacobian rows of the state, and the Newton matrix above kBlkRegMax variables.

struct BlkScratch {
  double *dphi, *gphi, *gdT, *gdm, *dH, *JC, *J, *M;
};
__host__ __device__ inline int64_t blk_scratch_doubles(const BlkLayout& L, int NS) {
  int64_t o = ((L.n_slots + 1) & ~1) + 4 * ((L.R + 1) & ~1) + static_cast<int64_t>(L.K) * L.K +
              ((static_cast<int64_t>(NS) * NS + 1) & ~int64_t(1));
  if (NS > kBlkRegMax) o += static_cast<int64_t>(NS) * kBlkThreads;
  return (o + 31) & ~int64_t(31);
}
HPO_DI BlkScratch make_blk_scratch(double* base, const BlkLayout& L) {
  BlkScratch s;
  double* p = base;
  s.dphi = p; p += (L.n_slots + 1) & ~1;
  s.gphi = p; p += (L.R + 1) & ~1;
  s.gdT = p; p += (L.R + 1) & ~1;
  s.gdm = p; p += (L.R + 1) & ~1;
  s.dH = p; p += (L.R + 1) & ~1;
  s.JC = p; p += static_cast<int64_t>(L.K) * L.K;
  s.J = p; p += (static_cast<int64_t>(L.K) * L.K + 1) & ~int64_t(1);
  s.M = p;   // only when K > kBlkRegMax
  return s;
}

// ---------------------------------------------------------------------------
// workgroup reductions: identical result in every thread
// ---------------------------------------------------------------------------
template <int N>
HPO_DI void bsum(double (&v)[N], double* red, int tid) {
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = wsum(v[j]);
  __syncthreads();   // the previous reduction's readers are done with red
  if ((tid & 63) == 0) {
#pragma unroll
    for (int j = 0; j < N; ++j) red[(tid >> 6) * 8 + j] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = (red[j] + red[8 + j]) + (red[16 + j] + red[24 + j]);
}
HPO_DI double bsum1(double v, double* red, int tid) {
  double a[1] = {v};
  bsum<1>(a, red, tid);
  return a[0];
}
HPO_DI double bmax(double v, double* red, int tid) {
  v = wmax(v);
  __syncthreads();
  if ((tid & 63) == 0) red[(tid >> 6) * 8] = v;
  __syncthreads();
  return fmax(fmax(red[0], red[8]), fmax(red[16], red[24]));
}
HPO_DI unsigned bmax_u(unsigned v, int* ired, int tid) {
  v = wmax_u(v);
  __syncthreads();
  if ((tid & 63) == 0) ired[tid >> 6] = static_cast<int>(v);
  __syncthreads();
  const unsigned a = static_cast<unsigned>(ired[0]), b = static_cast<unsigned>(ired[1]);
  const unsigned c = static_cast<unsigned>(ired[2]), d = static_cast<unsigned>(ired[3]);
  return max(max(a, b), max(c, d));
}

// ---------------------------------------------------------------------------
// species state on threads 0..K-1 (species k = thread k) from the broadcast state
// ---------------------------------------------------------------------------
HPO_DI void blk_species_state(const BlkTabs& t, const BlkLds& w, double P, int tid, SpState& s) {
  const int K = t.K;
  const double T = w.y[0];
  s.T = T;
  s.P = P;
  s.invT = 1.0 / T;
  s.lnT = fast_log(T);
  const int k = tid;
  double Yk = (k < K - 1) ? w.y[k + 1] : 0.0;
  const double* rec = t.sp + (k < K ? k : K - 1) * kSpStride;
  s.invW = rec[0];
  s.W = rec[1];
  double v[2] = {(k < K - 1) ? Yk : 0.0, (k < K - 1) ? Yk * s.invW : 0.0};
  bsum<2>(v, w.red, tid);
  const double ylast = 1.0 - v[0];
  if (k == K - 1) Yk = ylast;
  s.Yk = Yk;
  const double S = v[1] + ylast * t.sp[(K - 1) * kSpStride];
  s.Wbar = 1.0 / S;
  s.m = t.conv ? P * S : P * s.invT * (1.0 / kR);   // P / (R T), or rho / Wbar at constant volume
  s.rho = s.m * s.Wbar;
  s.invrho = S / s.m;
  s.lnPfac = kLnPatmOverR - s.lnT;
  const double* a = rec + 3 + (T <= rec[2] ? 0 : 7);
  s.cpR = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  s.hRT = a[0] + T * (a[1] * 0.5 + T * (a[2] * (1.0 / 3.0) + T * (a[3] * 0.25 + T * a[4] * 0.2))) + a[5] * s.invT;
  s.sR = a[0] * s.lnT + T * (a[1] + T * (a[2] * 0.5 + T * (a[3] * (1.0 / 3.0) + T * a[4] * 0.25))) + a[6];
  s.dcpR = a[1] + T * (2.0 * a[2] + T * (3.0 * a[3] + T * 4.0 * a[4]));
}

// Broadcasts the state, evaluates the species state and writes C, g, H/RT, cp (and the dummy
// species K) to LDS. Returns with those visible to the whole workgroup.
HPO_DI void blk_prepare(const BlkTabs& t, const BlkLds& w, double y, double P, int tid, int ns, SpState& s) {
  __syncthreads();   // previous readers of y / C / g are done
  w.y[tid] = (tid < ns) ? y : 0.0;
  __syncthreads();
  blk_species_state(t, w, P, tid, s);
  const int K = t.K;
  if (tid < K) {
    w.C[tid] = s.rho * s.Yk * s.invW;
    w.g[tid] = s.hRT - s.sR;
    w.hRT[tid] = s.hRT;
    w.cp[tid] = (s.cpR - t.conv) * kR * s.invW;   // mass cp, or cv at constant volume
  }
  if (tid == K) {
    w.C[K] = 1.0;
    w.g[K] = 0.0;
    w.hRT[K] = 0.0;
    w.cp[K] = 0.0;
  }
  __syncthreads();
}

// Third-body concentration and falloff of a pressure-dependent reaction (rxn_rates_pres_mod.c):
// pm, and d pm / d[M], d pm / dT at fixed [M] when DER.
template <bool DER>
HPO_DI void blk_pdep(const BlkTabs& t, const BlkLds& w, const SpState& s, const BlkRxn& h, double kf, double dlnkf,
                     double& pm, double& M, double& dpm_dM, double& dpm_dT) {
  const PdepRec& pr = t.pd[h.pidx];
  M = pr.eff_base * s.m;
  for (int e = pr.eff_begin; e < pr.eff_end; ++e) M += t.effv[e] * w.C[t.effs[e]];
  dpm_dM = 0.0;
  dpm_dT = 0.0;
  if (h.kind == kThreeBody) {
    pm = M;
    dpm_dM = 1.0;
    return;
  }
  const bool constk0 = h.flags & 8;
  const double k0 = arrh(pr.lnA0, pr.b0, pr.theta0, constk0, s.lnT, s.invT, t.exptab);
  const double dlnk0 = constk0 ? 0.0 : (pr.b0 + pr.theta0 * s.invT) * s.invT;
  const double Pr = k0 * M / kf;
  const double qq = 1.0 / (1.0 + Pr);
  if (h.kind == kTroe) {
    const double e3 = tab_exp(-s.T * pr.troe_iT3, t.exptab), e1 = tab_exp(-s.T * pr.troe_iT1, t.exptab);
    const double e2 = (h.flags & 4) ? tab_exp(-pr.troe_T2 * s.invT, t.exptab) : 0.0;
    const double Fc = (1.0 - pr.troe_a) * e3 + pr.troe_a * e1 + e2;
    const double logFc = fast_log10(fmax(Fc, 1.0e-300));
    const double x = fast_log10(fmax(Pr, 1.0e-300));
    const double A = x - 0.67 * logFc - 0.4;
    const double B = 0.806 - 1.1762 * logFc - 0.14 * x;
    const double D = 1.0 + A * A / (B * B);
    const double F = tab_exp(kLn10 * (logFc / D), t.exptab);
    pm = F * Pr * qq;
    if (DER) {
      const double dFc = -(1.0 - pr.troe_a) * pr.troe_iT3 * e3 - pr.troe_a * pr.troe_iT1 * e1 +
                         ((h.flags & 4) ? pr.troe_T2 * s.invT * s.invT * e2 : 0.0);
      const bool fcl = !(Fc > 1.0e-300), prl = !(Pr > 1.0e-300);
      const double dD_dA = 2.0 * A / (B * B), dD_dB = -2.0 * A * A / (B * B * B);
      const double df_dx = prl ? 0.0 : -logFc / (D * D) * (dD_dA - 0.14 * dD_dB);
      const double df_dc = 1.0 / D - logFc / (D * D) * (-0.67 * dD_dA - 1.1762 * dD_dB);
      const double dc_dT = fcl ? 0.0 : dFc / (Fc * kLn10);
      const double dF_dT = F * kLn10 * df_dc * dc_dT;
      const double dpm_dPr = F * qq * qq + F * df_dx * qq;
      dpm_dM = dpm_dPr * (k0 / kf);
      dpm_dT = dpm_dPr * Pr * (dlnk0 - dlnkf) + Pr * qq * dF_dT;
    }
  } else {
    pm = Pr * qq;
    if (DER) {
      dpm_dM = qq * qq * (k0 / kf);
      dpm_dT = qq * qq * Pr * (dlnk0 - dlnkf);
    }
  }
}

// Net rate (fwd - rev) * pm of reaction r.
HPO_DI double blk_rate(const BlkTabs& t, const BlkLds& w, const SpState& s, int r) {
  const BlkRxn h = t.rxn[r];
  const double* ar = t.arr + 3 * r;
  const double e = tab_exp(ar[0] + ar[1] * s.lnT - ar[2] * s.invT, t.exptab);
  const double kf = (h.flags & 2) ? ar[0] : e;
  double fwd = kf, pc = 1.0, dG = 0.0, dnu = 0.0;
#pragma unroll
  for (int q = 0; q < kBlkSide; ++q) {
    fwd *= cpow_int(w.C[h.sp[q]], h.nu[q]);
    dG -= h.nu[q] * w.g[h.sp[q]];
    dnu -= h.nu[q];
  }
#pragma unroll
  for (int q = kBlkSide; q < 2 * kBlkSide; ++q) {
    pc *= cpow_int(w.C[h.sp[q]], h.nu[q]);
    dG += h.nu[q] * w.g[h.sp[q]];
    dnu += h.nu[q];
  }
  const double rev = (h.flags & 1) ? kf * tab_exp(dG - dnu * s.lnPfac, t.exptab) * pc : 0.0;
  double pm = 1.0;
  if (h.kind != kElementary) {
    double M, a, b;
    blk_pdep<false>(t, w, s, h, kf, 0.0, pm, M, a, b);
  }
  return (fwd - rev) * pm;
}

// dydt (constant pressure, dydt.c:7-133) on the workgroup; thread t returns component t.
template <int NS>
HPO_DI double blk_rhs(const BlkTabs& t, const BlkLds& w, double y, double P, int tid) {
  const int K = t.K, R = t.R;
  SpState s;
  blk_prepare(t, w, y, P, tid, NS, s);
  for (int r = tid; r < R; r += kBlkThreads) w.phi[r] = blk_rate(t, w, s, r);
  __syncthreads();
  double wdot = 0.0;
  if (tid < K) {
    const int b = t.csr_ptr[tid], e = t.csr_ptr[tid + 1];
    for (int i = b; i < e; ++i) {
      const uint32_t c = t.csr[i];
      wdot += t.coef[c >> 16] * w.phi[c & 0xFFFFu];
    }
  }
  const bool sp = tid < K;
  // constant pressure: enthalpies and cp; constant volume: internal energies H - R T and cv
  const double cvf = t.conv;
  double v[2] = {sp ? (s.hRT - cvf) * kR * s.T * wdot : 0.0, sp ? s.Yk * (s.cpR - cvf) * kR * s.invW : 0.0};
  bsum<2>(v, w.red, tid);   // Q, cpmix
  // species row k lives in thread k + 1; the temperature row in thread 0
  if (tid < K - 1) w.vec[tid + 1] = wdot * s.W * s.invrho;
  if (tid == 0) w.vec[0] = -v[0] * s.invrho / v[1];
  __syncthreads();
  return tid < NS ? w.vec[tid] : 0.0;
}

// Analytic Jacobian (what jacob.c:3-371 computes): row t of J written to g.J[t * NS + j].
template <int NS>
HPO_DI void blk_jacobian(const BlkTabs& t, const BlkLds& w, const BlkScratch& g, double y, double P, int tid) {
  const int K = t.K, R = t.R;
  SpState s;
  blk_prepare(t, w, y, P, tid, NS, s);
  const double RT = kR * s.T;
  // per reaction: d(phi)/dC slots, phi, d(phi)/dT, d(phi)/dm, reaction enthalpy
  for (int r = tid; r < R; r += kBlkThreads) {
    const BlkRxn h = t.rxn[r];
    const double* ar = t.arr + 3 * r;
    const bool constk = h.flags & 2;
    const double kf = arrh(ar[0], ar[1], ar[2], constk, s.lnT, s.invT, t.exptab);
    const double dlnkf = constk ? 0.0 : (ar[1] + ar[2] * s.invT) * s.invT;
    double cr[kBlkSide], cq[kBlkSide];
    double fwd = kf;
#pragma unroll
    for (int q = 0; q < kBlkSide; ++q) {
      cr[q] = w.C[h.sp[q]];
      fwd *= cpow_int(cr[q], h.nu[q]);
    }
    double kr = 0.0, rev = 0.0, dlnkr = 0.0, dG = 0.0, dnu = 0.0, dHRT = 0.0;
#pragma unroll
    for (int q = 0; q < kBlkSide; ++q) {
      cq[q] = w.C[h.sp[kBlkSide + q]];
      const int n = h.nu[kBlkSide + q], sq = h.sp[kBlkSide + q];
      dG += n * w.g[sq];
      dHRT += n * w.hRT[sq];
      dnu += n;
    }
#pragma unroll
    for (int q = 0; q < kBlkSide; ++q) {
      const int n = h.nu[q], sq = h.sp[q];
      dG -= n * w.g[sq];
      dHRT -= n * w.hRT[sq];
      dnu -= n;
    }
    if (h.flags & 1) {
      kr = kf * tab_exp(dG - dnu * s.lnPfac, t.exptab);
      rev = kr;
#pragma unroll
      for (int q = 0; q < kBlkSide; ++q) rev *= cpow_int(cq[q], h.nu[kBlkSide + q]);
      dlnkr = dlnkf - (dHRT - dnu) * s.invT;
    }
    double pm = 1.0, M = 0.0, dpm_dM = 0.0, dpm_dT = 0.0, eff_base = 0.0;
    int eb = 0, ee = 0;
    if (h.kind != kElementary) {
      blk_pdep<true>(t, w, s, h, kf, dlnkf, pm, M, dpm_dM, dpm_dT);
      const PdepRec& pr = t.pd[h.pidx];
      eff_base = pr.eff_base;
      eb = pr.eff_begin;
      ee = pr.eff_end;
    }
    const double dq = fwd - rev;
    int sb = h.slot;
#pragma unroll
    for (int a = 0; a < kBlkSide; ++a)
      if (h.nu[a]) {
        double d = kf;
#pragma unroll
        for (int b = 0; b < kBlkSide; ++b)
          if (h.nu[b]) d *= (b == a) ? dcpow_int(cr[b], h.nu[b]) : cpow_int(cr[b], h.nu[b]);
        g.dphi[sb++] = pm * d;
      }
    if (h.flags & 1) {
#pragma unroll
      for (int a = 0; a < kBlkSide; ++a)
        if (h.nu[kBlkSide + a]) {
          double d = kr;
#pragma unroll
          for (int b = 0; b < kBlkSide; ++b)
            if (h.nu[kBlkSide + b])
              d *= (b == a) ? dcpow_int(cq[b], h.nu[kBlkSide + b]) : cpow_int(cq[b], h.nu[kBlkSide + b]);
          g.dphi[sb++] = -pm * d;
        }
    }
    const double fM = dq * dpm_dM;
    for (int e = eb; e < ee; ++e) g.dphi[sb++] = fM * t.effv[e];
    g.gphi[r] = dq * pm;
    g.gdm[r] = fM * eff_base;
    g.gdT[r] = pm * (fwd * dlnkf - rev * dlnkr) + dq * dpm_dT;
    if (h.flags & kBlkLumped) {
      // an irreversible reaction whose products are not in the slots: their enthalpy and moles
      for (int i = t.lump_ptr[r]; i < t.lump_ptr[r + 1]; ++i) {
        const uint32_t c = t.lump[i];
        const double nu = t.coef[c >> 16];
        dHRT += nu * w.hRT[c & 0xFFFFu];
        dnu += nu;
      }
    }
    g.dH[r] = (dHRT - t.conv * dnu) * RT;   // reaction enthalpy, or internal energy at constant volume
  }
  __syncthreads();
  // species rows (thread k): d(wdot_k)/dC row, wdot_k, d(wdot_k)/dT at fixed C and m,
  // d(wdot_k)/dm; and the enthalpy-weighted column sum of column k
  double* wdotv = w.sp4;
  double* JTv = w.sp4 + ((K + 1) & ~1);
  double* Jmv = w.sp4 + 2 * ((K + 1) & ~1);
  double* sigv = w.sp4 + 3 * ((K + 1) & ~1);
  double* dQv = w.sp4 + 4 * ((K + 1) & ~1);
  double wdot = 0.0, JT = 0.0, Jm = 0.0, sigma = 0.0, dQdC = 0.0;
  if (tid < K) {
    double* row = g.JC + static_cast<int64_t>(tid) * K;
    for (int i = 0; i < K; ++i) row[i] = 0.0;
    const int b = t.csr_ptr[tid], e = t.csr_ptr[tid + 1];
    for (int i = b; i < e; ++i) {
      const uint32_t c = t.csr[i];
      const int r = static_cast<int>(c & 0xFFFFu);
      const double nu = t.coef[c >> 16];
      wdot += nu * g.gphi[r];
      JT += nu * g.gdT[r];
      Jm += nu * g.gdm[r];
      const BlkRxn& h = t.rxn[r];
      for (int q = h.slot; q < h.slot + h.nslot; ++q) row[t.slot_sp[q]] += nu * g.dphi[q];
    }
    for (int i = 0; i < K; ++i) sigma += row[i] * w.C[i];
    const uint32_t cb = t.col_ptr[tid], ce = t.col_ptr[tid + 1];
    for (uint32_t i = cb; i < ce; ++i) {
      const uint32_t c = t.col[i];
      dQdC += g.dH[c >> 16] * g.dphi[c & 0xFFFFu];
    }
    wdotv[tid] = wdot;
    JTv[tid] = JT;
    Jmv[tid] = Jm;
    sigv[tid] = sigma;
    dQv[tid] = dQdC;
  }
  const bool sp = tid < K;
  const double cvf = t.conv, cpf = 1.0 - cvf;
  const double Hk = (s.hRT - cvf) * RT;   // enthalpy, or internal energy at constant volume
  double v[7] = {sp ? Hk * wdot : 0.0, sp ? (s.cpR - cvf) * kR * wdot : 0.0, sp ? Hk * JT : 0.0, sp ? Hk * Jm : 0.0,
                 sp ? s.Yk * (s.cpR - cvf) * kR * s.invW : 0.0, sp ? s.Yk * s.dcpR * kR * s.invW : 0.0,
                 sp ? dQdC * w.C[tid] : 0.0};
  bsum<7>(v, w.red, tid);   // also makes the species values above visible
  const double Q = v[0], SCpw = v[1], SHJT = v[2], SHJm = v[3], cpmix = v[4], dcpmix = v[5], sigT = v[6];
  // rows: thread 0 the temperature row, thread k + 1 the row of species k (chain rule to T, Y).
  // Constant pressure: C and m move with T (-C/T, -m/T) and C with Y through Wbar, and the
  // species rows' 1/rho moves too. Constant volume: C does not depend on T, m = rho sum Y/W
  // gives dm/dY_j = rho d_j, and rho is fixed.
  if (tid < NS) {
    const double rho = s.rho, Wbar = s.Wbar;
    const double invWN = t.sp[(K - 1) * kSpStride];
    const double D = rho * cpmix, invD = 1.0 / D, fT = -Q * invD;
    const double cpN = w.cp[K - 1];
    const double dcoef = -cpf * rho * Wbar * cpmix;
    double* Jrow = g.J + static_cast<int64_t>(tid) * NS;
    if (tid == 0) {
      const double dQdT = SCpw + SHJT - cpf * (sigT * s.invT + (s.m * s.invT) * SHJm);
      const double dDdT = -cpf * rho * cpmix * s.invT + rho * dcpmix;
      Jrow[0] = (-dQdT - fT * dDdT) * invD;
      const double c1 = dQv[K - 1] * rho * invWN;
      const double ysig = t.conv ? rho * SHJm : -Wbar * sigT;
      for (int j = 0; j < NS - 1; ++j) {
        const double invWj = t.sp[j * kSpStride];
        const double dj = invWj - invWN;
        const double base = dQv[j] * rho * invWj - c1 + dj * ysig;
        const double dDdY = dcoef * dj + rho * (w.cp[j] - cpN);
        Jrow[j + 1] = (-base - fT * dDdY) * invD;
      }
    } else {
      const int k = tid - 1;
      const double* row = g.JC + static_cast<int64_t>(k) * K;
      const double wk = wdotv[k], sk = sigv[k];
      const double fac = t.sp[k * kSpStride + 1] / rho;
      Jrow[0] = fac * (JTv[k] - cpf * (sk * s.invT + (s.m * s.invT) * Jmv[k] - wk * s.invT));
      const double c1 = row[K - 1] * rho * invWN;
      const double ysig = t.conv ? rho * Jmv[k] : -Wbar * sk;
      const double wcor = cpf * wk * Wbar;
      for (int j = 0; j < NS - 1; ++j) {
        const double invWj = t.sp[j * kSpStride];
        const double dj = invWj - invWN;
        const double base = row[j] * rho * invWj - c1 + dj * ysig;
        Jrow[j + 1] = fac * (base + wcor * dj);
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Newton matrix: Gauss-Jordan inverse across the workgroup, one register row per thread
// ---------------------------------------------------------------------------
template <int NS, typename F>
struct BlkInv {
  F a[NS];     // this thread's row; after the inversion, row `step` of (P M)^-1
  int step;    // the elimination step at which this row was the pivot (-1: none)
};

HPO_DI unsigned blk_pivot_key(double v, int tid) {
  const float f = fabsf(static_cast<float>(v));
  return (static_cast<unsigned>(__float_as_int(f)) & ~255u) | static_cast<unsigned>(255 - tid);
}
HPO_DI unsigned blk_pivot_key(float v, int tid) {
  return (static_cast<unsigned>(__float_as_int(fabsf(v))) & ~255u) | static_cast<unsigned>(255 - tid);
}

// Pivot: the largest |a[k]| among rows not yet pivoted (compared at single precision, the
// lower thread on ties). Every row, earlier pivot rows included, then subtracts l times the
// pivot row; column k takes the inverse's entries. Returns false for a singular matrix.
template <int NS, typename F>
HPO_DI bool blk_gj_invert_1b(BlkInv<NS, F>& L, const BlkLds& w, int tid) {
  // One barrier per step: each wave puts its best candidate row and its key into LDS, in one of
  // two buffers by the step's parity, and every thread then takes the row of the largest key.
  // A buffer is written again two steps later, after every thread has passed the barrier in
  // between, so after all its readers. (The same pivots and rows as a block-wide reduction
  // followed by the pivot thread's row, with two barriers fewer per step.)
  static_assert(kBlkWaves == 4, "four waves per workgroup");
  bool done = tid >= NS;
  bool singular = false;
  F rd = F(0);
  L.step = -1;
  const int wave = tid >> 6;
  F* cand = reinterpret_cast<F*>(w.prow);   // [2][4][NS]
  int* keys = w.iscal;                      // [2][4]
  __syncthreads();   // earlier users of these buffers are done
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int par = k & 1;
    const unsigned key = done ? 0u : blk_pivot_key(L.a[k], tid);
    const unsigned wk = wmax_u(key);
    if (tid == 255 - static_cast<int>(wk & 255u)) {
#pragma unroll
      for (int j = 0; j < NS; ++j) cand[(par * 4 + wave) * NS + j] = L.a[j];
    }
    if ((tid & 63) == 0) keys[par * 4 + wave] = static_cast<int>(wk);
    __syncthreads();
    const unsigned k0 = static_cast<unsigned>(keys[par * 4]), k1 = static_cast<unsigned>(keys[par * 4 + 1]);
    const unsigned k2 = static_cast<unsigned>(keys[par * 4 + 2]), k3 = static_cast<unsigned>(keys[par * 4 + 3]);
    const unsigned kmax = max(max(k0, k1), max(k2, k3));
    const int p = 255 - static_cast<int>(kmax & 255u);
    const F* prow = cand + (par * 4 + (p >> 6)) * NS;
    const bool isp = tid == p;
    if (isp) w.perm[k] = p;
    const F piv = prow[k];
    const bool bad = (kmax == 0u) || !(fabs(static_cast<double>(piv)) > 0.0);
    singular = singular || bad;
    F mult = bad ? F(0) : F(1) / piv;
    if (isp && !bad) {
      done = true;
      rd = mult;
      L.step = k;
    }
    const F l = isp ? F(0) : L.a[k] * mult;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      if (j == k) continue;
      L.a[j] = fma(-l, prow[j], L.a[j]);
    }
    L.a[k] = isp ? F(1) : -l;
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) L.a[j] = (tid < NS) ? L.a[j] * rd : F(0);
  __syncthreads();
  return !singular;
}

// The step-by-step form: a block-wide reduction for the pivot, then the pivot thread's row
// through LDS (three barriers per step). FP32 keeps it: with candidate rows per wave it ran
// 9 % slower on nDodecane.
template <int NS, typename F>
HPO_DI bool blk_gj_invert_2b(BlkInv<NS, F>& L, const BlkLds& w, int tid) {
  bool done = tid >= NS;
  bool singular = false;
  F rd = F(0);
  L.step = -1;
  F* prow = reinterpret_cast<F*>(w.prow);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const unsigned key = done ? 0u : blk_pivot_key(L.a[k], tid);
    const unsigned kmax = bmax_u(key, w.iscal, tid);
    const int p = 255 - static_cast<int>(kmax & 255u);
    const bool isp = tid == p;
    if (isp) {
#pragma unroll
      for (int j = 0; j < NS; ++j) prow[j] = L.a[j];
      w.perm[k] = p;
    }
    __syncthreads();
    const F piv = prow[k];
    const bool bad = (kmax == 0u) || !(fabs(static_cast<double>(piv)) > 0.0);
    singular = singular || bad;
    F mult = bad ? F(0) : F(1) / piv;
    if (isp && !bad) {
      done = true;
      rd = mult;
      L.step = k;
    }
    const F l = isp ? F(0) : L.a[k] * mult;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      if (j == k) continue;
      L.a[j] = fma(-l, prow[j], L.a[j]);
    }
    L.a[k] = isp ? F(1) : -l;
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) L.a[j] = (tid < NS) ? L.a[j] * rd : F(0);
  __syncthreads();
  return !singular;
}

// The inversion with the matrix in registers: the one-barrier form in FP64 (nDodecane 17 %
// faster), the step-by-step form in FP32. Both give the same result bit for bit.
template <int NS, typename F>
HPO_DI bool blk_gj_invert(BlkInv<NS, F>& L, const BlkLds& w, int tid) {
  if constexpr (sizeof(F) == 8)
    return blk_gj_invert_1b<NS, F>(L, w, tid);
  else
    return blk_gj_invert_2b<NS, F>(L, w, tid);
}

// x = M^-1 b (b, x: component t in thread t). Thread p_k holds row k of (P M)^-1, whose
// columns multiply b permuted to pivot order: b_{p_j} sits at position j, which thread p_j
// writes directly (its step is j). The product is x_k, sent to thread k.
template <int NS, typename F>
HPO_DI double blk_gj_apply(const BlkInv<NS, F>& L, double b, const BlkLds& w, int tid) {
  __syncthreads();   // previous readers of vec / vec2 are done
  // b in the matrix's precision, converted once by the thread that stores it rather than by
  // every reader (nDodecane FP32 -17 %, GRI block -7 %, same results; with the matrix in
  // scratch the same change made ESTiMatE 16 % slower, so blk_gj_apply_g keeps doubles)
  F* bv = reinterpret_cast<F*>(w.vec2);
  if (L.step >= 0) bv[L.step] = static_cast<F>(b);
  __syncthreads();
  F x0 = F(0), x1 = F(0);
#pragma unroll
  for (int j = 0; j + 1 < NS; j += 2) {
    x0 = fma(L.a[j], bv[j], x0);
    x1 = fma(L.a[j + 1], bv[j + 1], x1);
  }
  if (NS & 1) x0 = fma(L.a[NS - 1], bv[NS - 1], x0);
  if (L.step >= 0) w.vec[L.step] = static_cast<double>(x0 + x1);
  __syncthreads();
  return tid < NS ? w.vec[tid] : 0.0;
}

// The same elimination with the matrix in the workgroup's scratch (M[j * kBlkThreads + t] holds
// row t, column j; states of more than kBlkRegMax variables), in blocks of B = kBlkGjSteps
// steps, so that the matrix is read and written once per block rather than once per step:
// * step s of a block takes column k of its row through the block's earlier steps (its own
//   multipliers l[0..s) and the block's pivot rows in LDS), picks the pivot p, and builds
//   row p as it stands after those steps (one column per thread: row p's multipliers come
//   through LDS, and a column that was an earlier pivot column of the block starts from the
//   value row p took at that step);
// * then every column of the row takes the block's steps in order, one FMA each.
// Each entry sees the same FMAs in the same order as in the step-by-step elimination, so the
// result is the same bit for bit. On 512 ESTiMatE-Mech cells (214 variables) the blocks took
// the FP64 kernel from 3,412 to 986 ms and the FP32 one from 3,264 to 1,165 ms, with identical
// results; 4-step blocks in FP64 gave 1,353 ms and 16-step blocks in FP32 1,176 ms. Loading
// the sweep's columns (and blk_gj_apply_g's) in groups ahead of their updates then gave 908 /
// 1,073 ms. step: as BlkInv::step.
template <int NS, typename F>
HPO_DI bool blk_gj_invert_g(F* M, int& step, const BlkLds& w, int tid) {
  constexpr int B = kBlkGjSteps;
  bool done = tid >= NS;
  bool singular = false;
  F rd = F(0);
  step = -1;
  F* prow = reinterpret_cast<F*>(w.prow);   // [B][NS] the block's pivot rows
  F* lp = prow + B * NS;                    // [B] the current pivot row's multipliers
  for (int k0 = 0; k0 < NS; k0 += B) {
    const int nb = NS - k0 < B ? NS - k0 : B;
    F l[B];
#pragma unroll
    for (int s = 0; s < B; ++s) l[s] = F(0);
#pragma unroll
    for (int s = 0; s < B; ++s) {
      if (s >= nb) break;
      const int k = k0 + s;
      F akk = tid < NS ? M[k * kBlkThreads + tid] : F(0);
#pragma unroll
      for (int s2 = 0; s2 < s; ++s2) akk = fma(-l[s2], prow[s2 * NS + k], akk);
      const unsigned key = done ? 0u : blk_pivot_key(akk, tid);
      const unsigned kmax = bmax_u(key, w.iscal, tid);   // its barriers also end the last step's reads
      const int p = 255 - static_cast<int>(kmax & 255u);
      if (tid == p) {
#pragma unroll
        for (int s2 = 0; s2 < s; ++s2) lp[s2] = l[s2];
        w.perm[k] = p;
      }
      __syncthreads();
      if (tid < NS) {
        const int j = tid;
        F v;
        if (j >= k0 && j < k) {
          // an earlier pivot column of this block: row p took -l_p there (1 had it been that
          // step's pivot row, which only a singular step leaves undone)
          const int s3 = j - k0;
          v = p == w.perm[j] ? F(1) : -lp[s3];
          for (int s2 = s3 + 1; s2 < s; ++s2) v = fma(-lp[s2], prow[s2 * NS + j], v);
        } else {
          v = M[j * kBlkThreads + p];
#pragma unroll
          for (int s2 = 0; s2 < s; ++s2) v = fma(-lp[s2], prow[s2 * NS + j], v);
        }
        prow[s * NS + j] = v;
      }
      __syncthreads();
      const F piv = prow[s * NS + k];
      const bool bad = (kmax == 0u) || !(fabs(static_cast<double>(piv)) > 0.0);
      singular = singular || bad;
      const F mult = bad ? F(0) : F(1) / piv;
      const bool isp = tid == p;
      if (isp && !bad) {
        done = true;
        rd = mult;
        step = k;
      }
      l[s] = isp ? F(0) : akk * mult;
    }
    if (tid < NS) {
      constexpr int G = 8;   // groups of columns: their loads ahead of their updates
      for (int j0 = 0; j0 < NS; j0 += G) {
        F v[G];
#pragma unroll
        for (int q = 0; q < G; ++q) v[q] = j0 + q < NS ? M[(j0 + q) * kBlkThreads + tid] : F(0);
#pragma unroll
        for (int q = 0; q < G; ++q) {
#pragma unroll
          for (int s2 = 0; s2 < B; ++s2)
            if (s2 < nb && j0 + q < NS) v[q] = fma(-l[s2], prow[s2 * NS + j0 + q], v[q]);
        }
#pragma unroll
        for (int q = 0; q < G; ++q)
          if (j0 + q < NS) M[(j0 + q) * kBlkThreads + tid] = v[q];
      }
      // the block's own pivot columns (overwriting what the loop above stored there): the
      // value each row took at the column's step, then the block's later steps
#pragma unroll
      for (int s3 = 0; s3 < B; ++s3) {
        if (s3 >= nb) break;
        const int j = k0 + s3;
        F v = tid == w.perm[j] ? F(1) : -l[s3];
#pragma unroll
        for (int s2 = s3 + 1; s2 < B; ++s2)
          if (s2 < nb) v = fma(-l[s2], prow[s2 * NS + j], v);
        M[j * kBlkThreads + tid] = v;
      }
    }
    __syncthreads();   // the block's pivot rows are read; the next block writes them
  }
  if (tid < NS)
    for (int j = 0; j < NS; ++j) M[j * kBlkThreads + tid] *= rd;
  __syncthreads();
  return !singular;
}

template <int NS, typename F>
HPO_DI double blk_gj_apply_g(const F* M, int step, double b, const BlkLds& w, int tid) {
  __syncthreads();   // previous readers of vec / vec2 are done
  if (step >= 0) w.vec2[step] = b;
  __syncthreads();
  F x0 = F(0), x1 = F(0);
  if (tid < NS) {
    // groups of 16 column loads issued before their products (the same products in the same
    // order as one pair at a time)
    constexpr int G = 16;
    for (int j0 = 0; j0 < NS; j0 += G) {
      F m[G];
#pragma unroll
      for (int q = 0; q < G; ++q) m[q] = j0 + q < NS ? M[(j0 + q) * kBlkThreads + tid] : F(0);
#pragma unroll
      for (int q = 0; q < G; q += 2) {
        if (j0 + q + 1 < NS) {
          x0 = fma(m[q], static_cast<F>(w.vec2[j0 + q]), x0);
          x1 = fma(m[q + 1], static_cast<F>(w.vec2[j0 + q + 1]), x1);
        } else if (j0 + q < NS) {
          x0 = fma(m[q], static_cast<F>(w.vec2[j0 + q]), x0);
        }
      }
    }
  }
  if (step >= 0) w.vec[step] = static_cast<double>(x0 + x1);
  __syncthreads();
  return tid < NS ? w.vec[tid] : 0.0;
}

struct BlkInvStep {
  int step;
};

template <int NS, bool LU32>
struct BlockLayout {
  using F = typename std::conditional<LU32, float, double>::type;
  static constexpr bool kGlobalM = NS > kBlkRegMax;
  const BlkTabs& t;
  const BlkLds& w;
  const BlkScratch& g;
  const int tid;
  const int wave_lane;
  const double* exptab;
  typename std::conditional<kGlobalM, BlkInvStep, BlkInv<NS, F>>::type inv;

  HPO_DI BlockLayout(const BlkTabs& t_, const BlkLds& w_, const BlkScratch& g_, int tid_)
      : t(t_), w(w_), g(g_), tid(tid_), wave_lane(tid_ & 63), exptab(t_.exptab) {}
  HPO_DI bool any(bool c) const { return bsum1(c ? 1.0 : 0.0, w.red, tid) > 0.0; }
  HPO_DI double sum(double v) const { return bsum1(v, w.red, tid); }
  HPO_DI double max(double v) const { return bmax(v, w.red, tid); }
  HPO_DI double rhs(double y, double P) const { return blk_rhs<NS>(t, w, y, P, tid); }

  template <bool PROF>
  HPO_DI bool setup(double y, double P, double gamma, bool jnew, long long& cj) {
    if (jnew) {
      const long long c0 = PROF ? clock64() : 0;
      blk_jacobian<NS>(t, w, g, y, P, tid);
      if (PROF) cj = clock64() - c0;
    }
    const double* Jrow = g.J + static_cast<int64_t>(tid < NS ? tid : 0) * NS;
    if constexpr (kGlobalM) {
      F* M = reinterpret_cast<F*>(g.M);
      for (int j = 0; j < NS; ++j) {
        const double m = -gamma * Jrow[j] + ((j == tid) ? 1.0 : 0.0);
        M[j * kBlkThreads + tid] = tid < NS ? static_cast<F>(m) : F(0);
      }
      return blk_gj_invert_g<NS, F>(M, inv.step, w, tid);
    } else {
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const double m = -gamma * Jrow[j] + ((j == tid) ? 1.0 : 0.0);
        inv.a[j] = tid < NS ? static_cast<F>(m) : F(0);
      }
      return blk_gj_invert<NS, F>(inv, w, tid);
    }
  }

  HPO_DI double apply(double b) {
    if constexpr (kGlobalM)
      return blk_gj_apply_g<NS, F>(reinterpret_cast<const F*>(g.M), inv.step, b, w, tid);
    else
      return blk_gj_apply<NS, F>(inv, b, w, tid);
  }

  // x = M^-1 b; with the FP32 inverse, `sweeps` FP64 residual sweeps r = b - (I - gamma J) x
  // against the FP64 Jacobian rows in scratch.
  HPO_DI double solve(double b, double gl, int sweeps) {
    double x = apply(b);
    if constexpr (LU32) {
      const double* Jrow = g.J + static_cast<int64_t>(tid < NS ? tid : 0) * NS;
      for (int sw = 0; sw < sweeps; ++sw) {
        __syncthreads();
        w.y[tid] = x;
        __syncthreads();
        double Jx = 0.0;
        if constexpr (kGlobalM) {
          // groups of loads ahead of their products, in the same order (with the matrix in
          // registers the same grouping made the nDodecane and GRI kernels 75 % and 38 % slower)
          constexpr int G = 4;
          for (int j0 = 0; j0 < NS; j0 += G) {
            double v[G];
#pragma unroll
            for (int q = 0; q < G; ++q) v[q] = j0 + q < NS ? Jrow[j0 + q] : 0.0;
#pragma unroll
            for (int q = 0; q < G; ++q)
              if (j0 + q < NS) Jx = fma(v[q], w.y[j0 + q], Jx);
          }
        } else {
          for (int j = 0; j < NS; ++j) Jx = fma(Jrow[j], w.y[j], Jx);
        }
        const double r = (tid < NS) ? b - x + gl * Jx : 0.0;
        x += apply(r);
      }
    }
    return x;
  }
};

template <int NS, bool LU32>
constexpr int blk_waves_per_eu() {
  return LU32 ? HPO_BLK_WAVES_PER_EU_FP32 : HPO_BLK_WAVES_PER_EU_FP64;
}

template <int NS, bool LU32, bool PROF>
__global__ __launch_bounds__(kBlkThreads) __attribute__((
    amdgpu_waves_per_eu(blk_waves_per_eu<NS, LU32>(), blk_waves_per_eu<NS, LU32>()))) void bdf_block_kernel(BdfArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const BlkLayout& Lay = A.blay;
  for (int i = threadIdx.x * 16; i < Lay.lds_bytes; i += blockDim.x * 16)
    *reinterpret_cast<uint4*>(smem + i) = *reinterpret_cast<const uint4*>(A.mech + i);
  __syncthreads();
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  BlkTabs t = make_blk_tabs(smem, A.mech, Lay);
  t.conv = A.reactor;
  double* wbase = reinterpret_cast<double*>(smem + Lay.lds_bytes);
  BlkLds w = make_blk_lds(wbase, Lay.K, Lay.R);
  Scal& S = *reinterpret_cast<Scal*>(wbase + blk_lds_doubles(Lay.K, Lay.R) + wave * kScalDoubles);
  BlkScratch g = make_blk_scratch(A.scratch + static_cast<int64_t>(blockIdx.x) * A.scratch_stride, Lay);
  // An FP64 launch with the fallback armed takes only the cells that the FP32 launch before it
  // left failed, one per pull from queue[2], from their input rows (the words read alike in
  // every thread)
  if (!LU32 && hpo_fallback_armed(A.queue) && *reinterpret_cast<volatile uint32_t*>(A.queue + 1) == 0) return;
  for (;;) {
    const bool fb = !LU32 && hpo_fallback_armed(A.queue);
    __syncthreads();
    if (tid == 0) w.iscal[8] = static_cast<int>(atomicAdd(A.queue + (fb ? 2 : 0), 1u));
    __syncthreads();
    const unsigned c = static_cast<unsigned>(w.iscal[8]);
    if (static_cast<int64_t>(c) >= A.ncells) break;
    const int64_t cell = (!fb && A.order) ? static_cast<int64_t>(A.order[c]) : static_cast<int64_t>(c);
    if (fb && !hpo_fallback_status(A.status[cell])) continue;
    __syncthreads();
    blk_plog_cell(t, A.P[cell], tid);   // PLOG rates at this cell's pressure
    __syncthreads();
    BlockLayout<NS, LU32> Ly(t, w, g, tid);
    integrate_cell<NS, LU32, PROF>(A, Ly, S, cell);
  }
}

// ---- component kernels used by the parity tests ----
template <int NS, int MODE>
__global__ __launch_bounds__(kBlkThreads) void blk_eval_kernel(EvalArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const BlkLayout& Lay = A.blay;
  for (int i = threadIdx.x * 16; i < Lay.lds_bytes; i += blockDim.x * 16)
    *reinterpret_cast<uint4*>(smem + i) = *reinterpret_cast<const uint4*>(A.mech + i);
  __syncthreads();
  const int tid = threadIdx.x;
  BlkTabs t = make_blk_tabs(smem, A.mech, Lay);
  t.conv = A.reactor;
  BlkLds w = make_blk_lds(reinterpret_cast<double*>(smem + Lay.lds_bytes), Lay.K, Lay.R);
  BlkScratch g = make_blk_scratch(A.scratch + static_cast<int64_t>(blockIdx.x) * A.scratch_stride, Lay);
  for (int64_t c = blockIdx.x; c < A.ncells; c += gridDim.x) {
    __syncthreads();   // the previous cell's readers of the Arrhenius table are done
    blk_plog_cell(t, A.P[c], tid);
    __syncthreads();
    const double y = tid < NS ? A.y[c * A.ld + tid] : 0.0;
    const double P = A.P[c];
    if constexpr (MODE == 0) {
      const double f = blk_rhs<NS>(t, w, y, P, tid);
      if (tid < NS) A.out[c * A.ld + tid] = f;
    } else {
      blk_jacobian<NS>(t, w, g, y, P, tid);
      if (tid < NS)
        for (int j = 0; j < NS; ++j) A.out[(c * NS + tid) * NS + j] = g.J[tid * NS + j];
    }
  }
}

template <int NS, typename F>
__global__ __launch_bounds__(kBlkThreads) void blk_gj_kernel(LuArgs A) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int tid = threadIdx.x;
  BlkLds w = make_blk_lds(reinterpret_cast<double*>(smem), NS, 1);
  const int64_t c = blockIdx.x;
  BlkInv<NS, F> inv;
#pragma unroll
  for (int j = 0; j < NS; ++j) inv.a[j] = tid < NS ? static_cast<F>(A.M[(c * NS + tid) * NS + j]) : F(0);
  const bool ok = blk_gj_invert<NS, F>(inv, w, tid);
  const double b = tid < NS ? A.b[c * NS + tid] : 0.0;
  const double x = blk_gj_apply<NS, F>(inv, b, w, tid);
  if (tid < NS) A.x[c * NS + tid] = x;
  if (tid == 0) A.ok[c] = ok ? 1 : 0;
}

}  // namespace dev

// ---------------------------------------------------------------------------
// launch wrappers (host)
// ---------------------------------------------------------------------------
// State sizes with a block kernel: GRI-3.0 (53), nDodecane_Reitz (100), ESTiMatE-Mech (214, the
// Newton matrix in scratch). The Gauss-Jordan component kernel keeps its register row, so it
// exists for the first two only.
#define HPO_FOR_EACH_BLK_NS(X) X(53) X(100) X(214)
#define HPO_FOR_EACH_BLK_GJ_NS(X) X(53) X(100)

bool block_kernel_supports(int ns) {
#define HPO_CASE(N) \
  if (ns == N) return true;
  HPO_FOR_EACH_BLK_NS(HPO_CASE)
#undef HPO_CASE
  return false;
}

int blk_work_lds_bytes(const BlkLayout& L) { return blk_work_lds_bytes_for(L.K, L.R); }
int blk_work_lds_bytes_for(int K, int R) {
  return (dev::blk_lds_doubles(K, R) + dev::kBlkWaves * dev::kScalDoubles) * 8;
}
int64_t blk_scratch_doubles_host(const BlkLayout& L) { return dev::blk_scratch_doubles(L, L.K); }

template <int NS>
static const void* blk_fn(bool lu32, bool prof) {
  if (lu32) return prof ? reinterpret_cast<const void*>(&dev::bdf_block_kernel<NS, true, true>)
                        : reinterpret_cast<const void*>(&dev::bdf_block_kernel<NS, true, false>);
  return prof ? reinterpret_cast<const void*>(&dev::bdf_block_kernel<NS, false, true>)
              : reinterpret_cast<const void*>(&dev::bdf_block_kernel<NS, false, false>);
}

const void* bdf_block_kernel_fn(int ns, int lu32, int prof) {
#define HPO_CASE(N) \
  if (ns == N) return blk_fn<N>(lu32 != 0, prof != 0);
  HPO_FOR_EACH_BLK_NS(HPO_CASE)
#undef HPO_CASE
  return nullptr;
}

hipError_t launch_bdf_block(int ns, const BdfArgs& A, int lu32, int grid, size_t shmem, hipStream_t st) {
  const void* fn = bdf_block_kernel_fn(ns, lu32, A.prof != nullptr);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<BdfArgs*>(&A)};
  return hipLaunchKernel(fn, dim3(grid), dim3(dev::kBlkThreads), args, shmem, st);
}

hipError_t launch_eval_block(int ns, const EvalArgs& A, int grid, size_t shmem, hipStream_t st) {
#define HPO_CASE(N)                                                                                       \
  if (ns == N) {                                                                                          \
    if (A.mode == 0)                                                                                      \
      hipLaunchKernelGGL((dev::blk_eval_kernel<N, 0>), dim3(grid), dim3(dev::kBlkThreads), shmem, st, A); \
    else                                                                                                  \
      hipLaunchKernelGGL((dev::blk_eval_kernel<N, 1>), dim3(grid), dim3(dev::kBlkThreads), shmem, st, A); \
    return hipGetLastError();                                                                             \
  }
  HPO_FOR_EACH_BLK_NS(HPO_CASE)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_gj_block(int ns, int fp32, const LuArgs& A, int64_t ncases, hipStream_t st) {
  if (ncases <= 0) return hipSuccess;
  const size_t shmem = static_cast<size_t>(dev::blk_lds_doubles(ns, 1)) * 8;
#define HPO_CASE(N)                                                                                              \
  if (ns == N) {                                                                                                 \
    if (fp32)                                                                                                    \
      hipLaunchKernelGGL((dev::blk_gj_kernel<N, float>), dim3(ncases), dim3(dev::kBlkThreads), shmem, st, A);   \
    else                                                                                                         \
      hipLaunchKernelGGL((dev::blk_gj_kernel<N, double>), dim3(ncases), dim3(dev::kBlkThreads), shmem, st, A);  \
    return hipGetLastError();                                                                                    \
  }
  HPO_FOR_EACH_BLK_GJ_NS(HPO_CASE)
#undef HPO_CASE
  return hipErrorInvalidValue;
}

}  // namespace hpo
