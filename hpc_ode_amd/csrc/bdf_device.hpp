// Device building blocks for the per-cell BDF kernel on CDNA4 (gfx950).
//
// Layout: one wave64 integrates one cell. Lane l owns state component l
// (y = [T, Y_0 .. Y_{K-2}], as in src/Integrators/CVodeIntegrator.cpp:64-77), the
// matching row of the Newton matrix, and the matching entries of every BDF vector.
// The mechanism tables are staged once per workgroup into LDS; per-wave LDS holds the
// concentrations, Gibbs energies and reaction rates of the current RHS evaluation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fastmath.hpp"
#include "mechanism.hpp"

namespace hpo {
namespace dev {

#define HPO_DI __device__ __forceinline__

constexpr double kR = 8314.4621;
constexpr double kPatm = 101325.0;
constexpr double kLnPatmOverR = 2.5003367522806914;  // log(101325 / 8314.4621)

// ---------------------------------------------------------------------------
// cross-lane helpers (all lanes of the wave must be active)
// ---------------------------------------------------------------------------
HPO_DI double rl(double v, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
HPO_DI float rl(float v, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); }
HPO_DI int rli(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
// Lane `lane` of `old` takes the wave-uniform value v; other lanes keep theirs. A compare and
// select on the caller's (opaque) lane index rather than an inline-asm v_writelane: that
// instruction ignores EXEC, and as an asm operand the register allocator would not know it, so
// it could share the register with a value still live in lanes outside the current EXEC mask.
HPO_DI int wli(int old, int v, int lane, int my_lane) { return my_lane == lane ? v : old; }

template <int CTRL>
HPO_DI double dpp(double v) {
  int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
HPO_DI float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over all 64 lanes, identical result in every lane. quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror leave each row of 16 holding its
// sum; the four row sums are combined through readlane.
HPO_DI double wsum(double v) {
  v += dpp<0xB1>(v);
  v += dpp<0x4E>(v);
  v += dpp<0x141>(v);
  v += dpp<0x140>(v);
  return (rl(v, 0) + rl(v, 16)) + (rl(v, 32) + rl(v, 48));
}
// Two sums at once; the interleaved chains hide each other's DPP latency.
HPO_DI void wsum2(double& a, double& b) {
  a += dpp<0xB1>(a);
  b += dpp<0xB1>(b);
  a += dpp<0x4E>(a);
  b += dpp<0x4E>(b);
  a += dpp<0x141>(a);
  b += dpp<0x141>(b);
  a += dpp<0x140>(a);
  b += dpp<0x140>(b);
  a = (rl(a, 0) + rl(a, 16)) + (rl(a, 32) + rl(a, 48));
  b = (rl(b, 0) + rl(b, 16)) + (rl(b, 32) + rl(b, 48));
}
template <typename F>
HPO_DI F wmax(F v) {
  v = fmax(v, dpp<0xB1>(v));
  v = fmax(v, dpp<0x4E>(v));
  v = fmax(v, dpp<0x141>(v));
  v = fmax(v, dpp<0x140>(v));
  return fmax(fmax(rl(v, 0), rl(v, 16)), fmax(rl(v, 32), rl(v, 48)));
}

HPO_DI double shfl(double v, int src) { return __shfl(v, src, 64); }

// Returns v unchanged but hides its value from the optimiser. Used on the lane index where
// lane-dependent constants (identity-matrix entries, per-lane addresses) are formed, so
// loop-invariant code motion does not hoist dozens of them out of the cell loop and keep
// them live (and spilled) for the whole kernel.
HPO_DI int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Orders LDS and scratch accesses of the lanes of one wave (no workgroup barrier).
HPO_DI void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// Dynamic (wave-uniform) index into a small register array. Written as switches over
// constant indices so that the array stays promotable to registers: a loop of selects is
// folded back into an indexed load, which forces the array into scratch memory.
template <typename T>
HPO_DI T pick(const T (&v)[6], int i) {
  switch (i) {
    case 0: return v[0];
    case 1: return v[1];
    case 2: return v[2];
    case 3: return v[3];
    case 4: return v[4];
    default: return v[5];
  }
}
template <typename T>
HPO_DI void put(T (&v)[6], int i, T x) {
  switch (i) {
    case 0: v[0] = x; break;
    case 1: v[1] = x; break;
    case 2: v[2] = x; break;
    case 3: v[3] = x; break;
    case 4: v[4] = x; break;
    default: v[5] = x; break;
  }
}

// ---------------------------------------------------------------------------
// table views
// ---------------------------------------------------------------------------
struct Tabs {
  const double* sp;        // [K][kSpStride]
  const RxnHdr* hdr;       // [R]
  const RxnMol* mol;       // [R]
  const double* arr;       // [R][3]
  const PdepRec* pd;       // [n_pdep]
  const double* effv;      // [n_eff]
  const uint8_t* effs;     // [n_eff]
  const double* effpv;     // [n_pdep][neffp] padded efficiencies
  const uint8_t* effps;    // [n_pdep][neffp] their species
  const double* coef;      // [16]
  const uint32_t* gather;  // [ngl][64]
  const uint16_t* pptr;    // [K+1]
  const uint16_t* slot_base;  // [R+1] (global)
  const uint64_t* jterm;      // [njl][64] (global)
  const double* exptab;       // [32] 2^(j/32) for tab_exp
  int K, R, ngl, njl, npart, neffp, npdep, mol_ok;
  int conv;                   // reactor: 0 constant pressure (P is the pressure), 1 constant
                              // volume (P is the density)
};

HPO_DI Tabs make_tabs(const uint8_t* lds, const uint8_t* glob, const DevMechLayout& L) {
  Tabs t;
  t.sp = reinterpret_cast<const double*>(lds + L.off_species);
  t.hdr = reinterpret_cast<const RxnHdr*>(lds + L.off_hdr);
  t.mol = reinterpret_cast<const RxnMol*>(lds + L.off_mol);
  t.mol_ok = L.mol_ok;
  t.arr = reinterpret_cast<const double*>(lds + L.off_arr);
  t.pd = reinterpret_cast<const PdepRec*>(lds + L.off_pdep);
  t.effv = reinterpret_cast<const double*>(lds + L.off_eff_val);
  t.effs = lds + L.off_eff_sp;
  t.effpv = reinterpret_cast<const double*>(lds + L.off_effp_val);
  t.effps = lds + L.off_effp_sp;
  t.neffp = L.n_eff_pad;
  t.npdep = L.n_pdep;
  t.coef = reinterpret_cast<const double*>(lds + L.off_coef);
  t.gather = reinterpret_cast<const uint32_t*>(lds + L.off_gather);
  t.pptr = reinterpret_cast<const uint16_t*>(lds + L.off_part_ptr);
  t.exptab = reinterpret_cast<const double*>(lds + L.off_exptab);
  t.slot_base = reinterpret_cast<const uint16_t*>(glob + L.off_slot_base);
  t.jterm = reinterpret_cast<const uint64_t*>(glob + L.off_jterm);
  t.K = L.K;
  t.R = L.R;
  t.ngl = L.n_gather_lane;
  t.njl = L.n_jterm_lane;
  t.npart = L.n_partials;
  t.conv = 0;
  return t;
}

// per-wave LDS: C[64], g[64], hRT[64], cp[64] (mass), phi[R], part[3*npart]
struct WaveLds {
  double *C, *g, *hRT, *cp, *phi, *part;
};
__host__ __device__ inline int wave_lds_doubles(int R, int npart) { return 256 + ((R + 1) & ~1) + 3 * ((npart + 2) & ~1); }
HPO_DI WaveLds make_wave_lds(double* base, int R, int npart) {
  WaveLds w;
  w.C = base;
  w.g = base + 64;
  w.hRT = base + 128;
  w.cp = base + 192;
  w.phi = base + 256;
  w.part = base + 256 + ((R + 1) & ~1);
  return w;
}

// per-wave global scratch (doubles)
struct WaveScratch {
  double *dphi, *coefbuf, *gphi, *gdm, *gdT, *JC, *J;
};
HPO_DI WaveScratch make_wave_scratch(double* base, const DevMechLayout& L) {
  WaveScratch s;
  int o = 0;
  s.JC = base + o;
  o += 64 * 64;
  s.J = base + o;
  o += 64 * 64;
  s.dphi = base + o;
  o += (L.n_slots + 1) & ~1;
  s.coefbuf = base + o;
  o += 16 + ((L.R + 1) & ~1);
  s.gphi = base + o;
  o += (L.R + 1) & ~1;
  s.gdm = base + o;
  o += (L.R + 1) & ~1;
  s.gdT = base + o;
  return s;
}
inline int64_t wave_scratch_doubles(const DevMechLayout& L) {
  int64_t o = 2 * 64 * 64 + ((L.n_slots + 1) & ~1) + 16 + 4 * ((L.R + 1) & ~1);
  return (o + 31) & ~int64_t(31);
}

// ---------------------------------------------------------------------------
// species-level state shared by the RHS and the Jacobian
// ---------------------------------------------------------------------------
struct SpState {
  double T, lnT, invT, P, rho, invrho, Wbar, m, lnPfac;
  double Yk, invW, W, cpR, hRT, sR, dcpR;  // per lane (species k = lane)
};

HPO_DI void species_state(const Tabs& t, double y, double P, int lane_in, SpState& s) {
  const int lane = opaque(lane_in);
  const int K = t.K;
  const double T = rl(y, 0);
  s.T = T;
  s.P = P;
  s.invT = 1.0 / T;
  s.lnT = fast_log(T);
  double Yk = shfl(y, (lane + 1) & 63);
  const double* rec = t.sp + (lane < K ? lane : K - 1) * kSpStride;
  s.invW = rec[0];
  s.W = rec[1];
  // sum Y_k (k < K-1) and sum Y_k/W_k (k < K-1) in one pair of interleaved reductions; the
  // eliminated species adds (1 - sum Y) / W_{K-1}
  double ys = (lane < K - 1) ? Yk : 0.0;
  double yw = (lane < K - 1) ? Yk * s.invW : 0.0;
  wsum2(ys, yw);
  const double ylast = 1.0 - ys;
  if (lane == K - 1) Yk = ylast;
  if (lane >= K) Yk = 0.0;
  s.Yk = Yk;
  double S = yw + ylast * t.sp[(K - 1) * kSpStride];
  s.Wbar = 1.0 / S;
  // total concentration: P / (R T) at constant pressure; at constant volume P holds the
  // density and m = rho / Wbar
  s.m = t.conv ? P * S : P * s.invT * (1.0 / kR);
  s.rho = s.m * s.Wbar;
  s.invrho = S / s.m;
  s.lnPfac = kLnPatmOverR - s.lnT;   // log(P_atm / (R T))
  const double* a = rec + 3 + (T <= rec[2] ? 0 : 7);
  s.cpR = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
  s.hRT = a[0] + T * (a[1] * 0.5 + T * (a[2] * (1.0 / 3.0) + T * (a[3] * 0.25 + T * a[4] * 0.2))) + a[5] * s.invT;
  s.sR = a[0] * s.lnT + T * (a[1] + T * (a[2] * 0.5 + T * (a[3] * (1.0 / 3.0) + T * a[4] * 0.25))) + a[6];
  s.dcpR = a[1] + T * (2.0 * a[2] + T * (3.0 * a[3] + T * 4.0 * a[4]));
}

// c^nu for nu in 0..3 without branches (nu = 0 marks an unused reaction slot)
HPO_DI double cpow_int(double c, int nu) {
  double c2 = c * c;
  double p = (nu >= 2) ? c2 : c;
  p = (nu >= 3) ? p * c : p;
  return (nu == 0) ? 1.0 : p;
}
HPO_DI double dcpow_int(double c, int nu) { return nu == 1 ? 1.0 : (nu == 2 ? 2.0 * c : 3.0 * c * c); }

HPO_DI double arrh(double lnA, double b, double th, bool constk, double lnT, double invT, const double* et) {
  return constk ? lnA : tab_exp(lnA + b * lnT - th * invT, et);
}
constexpr double kLn10 = 2.302585092994045684;

// Balanced gather: out[k] = sum_r nu_kr src[r] for species k = lane (0 for lane >= K).
// Each lane walks its share of the nonzeros, flushing partial sums per species run;
// lane k then adds the partials of species k.
template <int NSRC>
HPO_DI void gather(const Tabs& t, const double* const (&src)[NSRC], double* part, int lane_in, double (&out)[NSRC]) {
  const int lane = opaque(lane_in);
  double acc[NSRC];
#pragma unroll
  for (int s = 0; s < NSRC; ++s) acc[s] = 0.0;
  // Entries are read in chunks: the table words and source values of a chunk are all loaded
  // before the chunk's partial sums are stored, so the loads are not serialised behind stores
  // the compiler cannot prove independent. Entries past the table end read as padding
  // (reaction 0, coefficient 0, no flush).
  constexpr int CH = 4;
#pragma unroll 1
  for (int i0 = 0; i0 < t.ngl; i0 += CH) {
    uint32_t ev[CH];
    double v[CH][NSRC];
#pragma unroll
    for (int u = 0; u < CH; ++u) ev[u] = (i0 + u < t.ngl) ? t.gather[(i0 + u) * 64 + lane] : 0u;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int r = static_cast<int>(ev[u] & 4095u);
#pragma unroll
      for (int s = 0; s < NSRC; ++s) v[u][s] = src[s][r];
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const uint32_t e = ev[u];
      const double c = static_cast<double>(static_cast<int32_t>(e << 16) >> 28);
      // flush the partial sum at the end of a species run; other entries write a scratch slot
      const bool flush = (e >> 16) & 1u;
      const int p = flush ? static_cast<int>(e >> 17) : t.npart;
#pragma unroll
      for (int s = 0; s < NSRC; ++s) {
        acc[s] += c * v[u][s];
        part[s * (t.npart + 1) + p] = acc[s];
        acc[s] = flush ? 0.0 : acc[s];
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int s = 0; s < NSRC; ++s) out[s] = 0.0;
  if (lane < t.K) {
    int b = t.pptr[lane], e = t.pptr[lane + 1];
    for (int p = b; p < e; ++p) {
#pragma unroll
      for (int s = 0; s < NSRC; ++s) out[s] += part[s * (t.npart + 1) + p];
    }
  }
  wave_sync();
}

// Net rate (fwd - rev) * pressure modification of reaction r (rxn_rates.c, rxn_rates_pres_mod.c).
template <bool PDEP, bool MOL>
HPO_DI double reaction_rate(const Tabs& t, const WaveLds& w, const SpState& s, int r) {
  const RxnHdr h = t.hdr[r];
  const double* ar = t.arr + 3 * r;
  // straight-line for the elementary part, so several reactions per lane can be interleaved
  double e = tab_exp(ar[0] + ar[1] * s.lnT - ar[2] * s.invT, t.exptab);
  double kf = (h.flags & 2) ? ar[0] : e;
  double fwd = kf, pc = 1.0, dG = 0.0, dnu = 0.0;
  if constexpr (MOL) {
    // molecule slots: unused ones read the dummy species (C = 1, g = 0)
    const RxnMol mo = t.mol[r];
    fwd *= w.C[mo.m[0]] * w.C[mo.m[1]] * w.C[mo.m[2]];
    pc = w.C[mo.m[3]] * w.C[mo.m[4]] * w.C[mo.m[5]];
    dG = (w.g[mo.m[3]] + w.g[mo.m[4]] + w.g[mo.m[5]]) - (w.g[mo.m[0]] + w.g[mo.m[1]] + w.g[mo.m[2]]);
    dnu = static_cast<double>(mo.dnu);
  } else {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      fwd *= cpow_int(w.C[h.sp[q]], h.nu[q]);
      dG -= h.nu[q] * w.g[h.sp[q]];
      dnu -= h.nu[q];
    }
#pragma unroll
    for (int q = 3; q < 6; ++q) {
      pc *= cpow_int(w.C[h.sp[q]], h.nu[q]);
      dG += h.nu[q] * w.g[h.sp[q]];
      dnu += h.nu[q];
    }
  }
  // k_r = k_f / Kc, Kc = (P_atm/(R T))^dnu exp(-dG)
  double rev = (h.flags & 1) ? kf * tab_exp(dG - dnu * s.lnPfac, t.exptab) * pc : 0.0;
  double pm = 1.0;
  if (PDEP && h.kind != kElementary) {
    const PdepRec& pr = t.pd[h.pidx];
    double M = pr.eff_base * s.m;
    // fixed-length padded list (padding: species with C = 1, efficiency 0): all loads of a
    // group of four are independent, and the sum keeps the order of the real entries
    const double* ev = t.effpv + h.pidx * t.neffp;
    const uint8_t* es = t.effps + h.pidx * t.neffp;
    for (int e2 = 0; e2 < t.neffp; e2 += 4) {
      const uint32_t sp4 = *reinterpret_cast<const uint32_t*>(es + e2);
      const double c0 = w.C[sp4 & 255u], c1 = w.C[(sp4 >> 8) & 255u];
      const double c2 = w.C[(sp4 >> 16) & 255u], c3 = w.C[sp4 >> 24];
      M += ev[e2] * c0;
      M += ev[e2 + 1] * c1;
      M += ev[e2 + 2] * c2;
      M += ev[e2 + 3] * c3;
    }
    if (h.kind == kThreeBody) {
      pm = M;
    } else {
      double k0 = arrh(pr.lnA0, pr.b0, pr.theta0, h.flags & 8, s.lnT, s.invT, t.exptab);
      double Pr = k0 * M / kf;
      double F = 1.0;
      if (h.kind == kTroe) {
        double Fc = (1.0 - pr.troe_a) * tab_exp(-s.T * pr.troe_iT3, t.exptab) +
                    pr.troe_a * tab_exp(-s.T * pr.troe_iT1, t.exptab) +
                    ((h.flags & 4) ? tab_exp(-pr.troe_T2 * s.invT, t.exptab) : 0.0);
        double logFc = fast_log10(fmax(Fc, 1.0e-300));
        double x = fast_log10(fmax(Pr, 1.0e-300));
        double A = x - 0.67 * logFc - 0.4;
        double B = 0.806 - 1.1762 * logFc - 0.14 * x;
        const double B2 = B * B;
        F = tab_exp(kLn10 * (logFc * B2 / (B2 + A * A)), t.exptab);   // 10^(logFc / (1 + A^2/B^2))
      }
      pm = F * Pr / (1.0 + Pr);
    }
  }
  return (fwd - rev) * pm;
}

// ---------------------------------------------------------------------------
// RHS: constant-pressure dydt (src/Mechanism/CPU/dydt.c:7-133 in table form)
// ---------------------------------------------------------------------------
HPO_DI double rhs(const Tabs& t, const WaveLds& w, double y, double P, int lane_in, int n) {
  // lane-derived addresses and masks are formed per call (see opaque)
  const int lane = opaque(lane_in);
  const int K = t.K, R = t.R;
  SpState s;
  species_state(t, y, P, lane, s);
  if (lane < K) {
    w.C[lane] = s.rho * s.Yk * s.invW;
    w.g[lane] = s.hRT - s.sR;
  }
  if (lane == kDummySpecies) {
    w.C[lane] = 1.0;
    w.g[lane] = 0.0;
  }
  wave_sync();
  // Four reactions per lane per pass: independent exp chains the scheduler interleaves.
  // Pressure-dependent reactions come first in the device order; with at most 64 of them
  // only the first slot of the first pass can hold one, and the other slots take the
  // elementary-only path (a mechanism with more takes the general path in every slot).
  // The fast path (molecule slots, at most 64 pressure-dependent reactions) covers the
  // shipped mechanisms; any other takes the general path in every slot.
  const bool fast = t.npdep <= 64 && t.mol_ok;
#pragma unroll 1
  for (int base = 0; base < R; base += 256) {
    const int r0 = base + lane, r1 = r0 + 64, r2 = r0 + 128, r3 = r0 + 192;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    if (fast) {
      p0 = (r0 < R) ? reaction_rate<true, true>(t, w, s, r0) : 0.0;
      p1 = (r1 < R) ? reaction_rate<false, true>(t, w, s, r1) : 0.0;
      p2 = (r2 < R) ? reaction_rate<false, true>(t, w, s, r2) : 0.0;
      p3 = (r3 < R) ? reaction_rate<false, true>(t, w, s, r3) : 0.0;
    } else {
      p0 = (r0 < R) ? reaction_rate<true, false>(t, w, s, r0) : 0.0;
      p1 = (r1 < R) ? reaction_rate<true, false>(t, w, s, r1) : 0.0;
      p2 = (r2 < R) ? reaction_rate<true, false>(t, w, s, r2) : 0.0;
      p3 = (r3 < R) ? reaction_rate<true, false>(t, w, s, r3) : 0.0;
    }
    if (r0 < R) w.phi[r0] = p0;
    if (r1 < R) w.phi[r1] = p1;
    if (r2 < R) w.phi[r2] = p2;
    if (r3 < R) w.phi[r3] = p3;
  }
  wave_sync();
  const double* src[1] = {w.phi};
  double wd[1];
  gather<1>(t, src, w.part, lane, wd);
  const double wdot = wd[0];
  // constant pressure: molar enthalpy and cp; constant volume: internal energy H - R T and cv
  const double cvf = t.conv;
  const double Hk = (s.hRT - cvf) * kR * s.T;
  double Q = wsum(lane < K ? Hk * wdot : 0.0);
  double cpmix = wsum(lane < K ? s.Yk * (s.cpR - cvf) * kR * s.invW : 0.0);
  double fY = wdot * s.W * s.invrho;
  double f = shfl(fY, (lane + 63) & 63);
  if (lane == 0) f = -Q * s.invrho / cpmix;
  if (lane >= n) f = 0.0;
  return f;
}

// ---------------------------------------------------------------------------
// Analytic Jacobian (what src/Mechanism/CPU/jacob.c:3-371 computes), one row per lane.
// ---------------------------------------------------------------------------
// The row of lane l < NS is written to g.J[j*NS + l], j = 0..NS-1 (columns packed at the
// state length, so a column read moves NS doubles, not 64).
template <int NS>
HPO_DI void jacobian(const Tabs& t, const WaveLds& w, const WaveScratch& g, double y, double P, int lane_in) {
  const int lane = opaque(lane_in);
  const int K = t.K, R = t.R;
  SpState s;
  species_state(t, y, P, lane, s);
  // constant volume replaces enthalpies by internal energies (H - R T) and cp by cv
  const double cvf = t.conv;
  const double cpk = (s.cpR - cvf) * kR * s.invW;  // mass cp (cv) of species k
  if (lane < K) {
    w.C[lane] = s.rho * s.Yk * s.invW;
    w.g[lane] = s.hRT - s.sR;
    w.hRT[lane] = s.hRT;
    w.cp[lane] = cpk;
  }
  if (lane == kDummySpecies) {
    w.C[lane] = 1.0;
    w.g[lane] = 0.0;
    w.hRT[lane] = 0.0;
  }
  // zero the dense d wdot/dC buffer (columns 0..K-1, rows 0..63)
  for (int i = 0; i < K; ++i) g.JC[i * 64 + lane] = 0.0;
  wave_sync();
  const double RT = kR * s.T;
#pragma unroll 1
  for (int r = lane; r < R; r += 64) {
    const RxnHdr h = t.hdr[r];
    const double* ar = t.arr + 3 * r;
    const bool constk = h.flags & 2;
    double kf = arrh(ar[0], ar[1], ar[2], constk, s.lnT, s.invT, t.exptab);
    double dlnkf = constk ? 0.0 : (ar[1] + ar[2] * s.invT) * s.invT;
    double cr[3], cp_[3];
    double fwd = kf;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      cr[q] = w.C[h.sp[q]];
      fwd *= cpow_int(cr[q], h.nu[q]);
    }
    double kr = 0.0, rev = 0.0, dlnkr = 0.0, dH = 0.0;
    {
      double dG = 0.0, dnu = 0.0, dHRT = 0.0;
#pragma unroll
      for (int q = 3; q < 6; ++q) {
        cp_[q - 3] = w.C[h.sp[q]];
        if (h.nu[q]) {
          dG += h.nu[q] * w.g[h.sp[q]];
          dHRT += h.nu[q] * w.hRT[h.sp[q]];
          dnu += h.nu[q];
        }
      }
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (h.nu[q]) {
          dG -= h.nu[q] * w.g[h.sp[q]];
          dHRT -= h.nu[q] * w.hRT[h.sp[q]];
          dnu -= h.nu[q];
        }
      dH = (dHRT - cvf * dnu) * RT;
      if (h.flags & 1) {
        kr = kf * tab_exp(dG - dnu * s.lnPfac, t.exptab);
        rev = kr;
#pragma unroll
        for (int q = 0; q < 3; ++q)
          rev *= cpow_int(cp_[q], h.nu[3 + q]);
        dlnkr = dlnkf - (dHRT - dnu) * s.invT;
      }
    }
    double pm = 1.0, dpm_dM = 0.0, dpm_dT = 0.0, eff_base = 0.0;
    int eb = 0, ee = 0;
    if (h.kind != kElementary) {
      const PdepRec& pr = t.pd[h.pidx];
      eff_base = pr.eff_base;
      eb = pr.eff_begin;
      ee = pr.eff_end;
      double M = pr.eff_base * s.m;
      for (int e = eb; e < ee; ++e) M += t.effv[e] * w.C[t.effs[e]];
      if (h.kind == kThreeBody) {
        pm = M;
        dpm_dM = 1.0;
      } else {
        const bool constk0 = h.flags & 8;
        double k0 = arrh(pr.lnA0, pr.b0, pr.theta0, constk0, s.lnT, s.invT, t.exptab);
        double dlnk0 = constk0 ? 0.0 : (pr.b0 + pr.theta0 * s.invT) * s.invT;
        double Pr = k0 * M / kf;
        double dPr_dT = Pr * (dlnk0 - dlnkf);
        double dPr_dM = k0 / kf;
        double qq = 1.0 / (1.0 + Pr);
        if (h.kind == kTroe) {
          double e3 = tab_exp(-s.T * pr.troe_iT3, t.exptab), e1 = tab_exp(-s.T * pr.troe_iT1, t.exptab);
          double e2 = (h.flags & 4) ? tab_exp(-pr.troe_T2 * s.invT, t.exptab) : 0.0;
          double Fc = (1.0 - pr.troe_a) * e3 + pr.troe_a * e1 + e2;
          double dFc = -(1.0 - pr.troe_a) * pr.troe_iT3 * e3 - pr.troe_a * pr.troe_iT1 * e1 +
                       ((h.flags & 4) ? pr.troe_T2 * s.invT * s.invT * e2 : 0.0);
          bool fcl = !(Fc > 1.0e-300), prl = !(Pr > 1.0e-300);
          double logFc = fast_log10(fmax(Fc, 1.0e-300));
          double x = fast_log10(fmax(Pr, 1.0e-300));
          double A = x - 0.67 * logFc - 0.4;
          double B = 0.806 - 1.1762 * logFc - 0.14 * x;
          double D = 1.0 + A * A / (B * B);
          double F = tab_exp(kLn10 * (logFc / D), t.exptab);
          double dD_dA = 2.0 * A / (B * B), dD_dB = -2.0 * A * A / (B * B * B);
          double df_dx = prl ? 0.0 : -logFc / (D * D) * (dD_dA - 0.14 * dD_dB);
          double df_dc = 1.0 / D - logFc / (D * D) * (-0.67 * dD_dA - 1.1762 * dD_dB);
          double dc_dT = fcl ? 0.0 : dFc / (Fc * 2.302585092994046);
          double dF_dT = F * 2.302585092994046 * df_dc * dc_dT;
          pm = F * Pr * qq;
          double dpm_dPr = F * qq * qq + F * df_dx * qq;
          dpm_dM = dpm_dPr * dPr_dM;
          dpm_dT = dpm_dPr * dPr_dT + Pr * qq * dF_dT;
        } else {
          pm = Pr * qq;
          dpm_dM = qq * qq * dPr_dM;
          dpm_dT = qq * qq * dPr_dT;
        }
      }
    }
    const double dq = fwd - rev;
    int sb = t.slot_base[r];
    // reactant slots: pm * d fwd / dC
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (h.nu[a]) {
        double d = kf;
#pragma unroll
        for (int b = 0; b < 3; ++b)
          if (h.nu[b]) d *= (b == a) ? dcpow_int(cr[b], h.nu[b]) : cpow_int(cr[b], h.nu[b]);
        g.dphi[sb++] = pm * d;
      }
    if (h.flags & 1) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (h.nu[3 + a]) {
          double d = kr;
#pragma unroll
          for (int b = 0; b < 3; ++b)
            if (h.nu[3 + b]) d *= (b == a) ? dcpow_int(cp_[b], h.nu[3 + b]) : cpow_int(cp_[b], h.nu[3 + b]);
          g.dphi[sb++] = -pm * d;
        }
    }
    const double fM = dq * dpm_dM;
    for (int e = eb; e < ee; ++e) g.dphi[sb++] = fM * t.effv[e];
    g.gphi[r] = dq * pm;
    g.gdm[r] = fM * eff_base;
    g.gdT[r] = pm * (fwd * dlnkf - rev * dlnkr) + dq * dpm_dT;
    g.coefbuf[16 + r] = dH;
  }
  wave_sync();
  // owner-computes assembly of d wdot / dC (rows 0..K-2) and sum_k H_k d wdot_k / dC (row K)
  {
    // chunks of terms: all loads of a chunk (term words, then coefficient and slot values)
    // are issued before its stores. Terms past the table end read as coefficient 0, slot 0,
    // no flush; every lane's list ends with a flushed term, so they are never stored.
    constexpr int CH = 8;
    double acc = 0.0;
#pragma unroll 1
    for (int i0 = 0; i0 < t.njl; i0 += CH) {
      uint64_t ev[CH];
      double cv[CH], dv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) ev[u] = (i0 + u < t.njl) ? t.jterm[(i0 + u) * 64 + lane] : 0ull;
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        cv[u] = g.coefbuf[static_cast<int>((ev[u] >> 32) & 0xFFFFu)];
        dv[u] = g.dphi[static_cast<int>((ev[u] >> 16) & 0xFFFFu)];
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        acc += cv[u] * dv[u];
        if ((ev[u] >> 48) & 1u) {
          g.JC[static_cast<int>(ev[u] & 0xFFFFu)] = acc;
          acc = 0.0;
        }
      }
    }
  }
  // species gathers of wdot, d wdot/dT|C,m and d wdot/dm
  const double* srcs[3] = {g.gphi, g.gdT, g.gdm};
  double gv[3];
  gather<3>(t, srcs, w.part, lane, gv);  // contains wave_sync
  const double wdot = gv[0], JT = gv[1], Jm = gv[2];
  const double Hk = (s.hRT - cvf) * RT;
  const bool sp = lane < K;
  const double Q = wsum(sp ? Hk * wdot : 0.0);
  const double SCpw = wsum(sp ? (s.cpR - cvf) * kR * wdot : 0.0);
  const double SHJT = wsum(sp ? Hk * JT : 0.0);
  const double SHJm = wsum(sp ? Hk * Jm : 0.0);
  const double cpmix = wsum(sp ? s.Yk * cpk : 0.0);
  const double dcpmix = wsum(sp ? s.Yk * s.dcpR * kR * s.invW : 0.0);
  // species row k = lane-1 lives in lane k+1
  const double wdot_k = shfl(wdot, (lane + 63) & 63);
  const double JT_k = shfl(JT, (lane + 63) & 63);
  const double Jm_k = shfl(Jm, (lane + 63) & 63);
  // this lane's row of the buffer: species row lane-1, or the enthalpy-weighted row K
  const double* jrow = g.JC + ((lane == 0) ? K : ((lane <= K) ? lane - 1 : 63));
  double sigma = 0.0;
#pragma unroll 4
  for (int i = 0; i < NS; ++i) sigma += jrow[i * 64] * w.C[i];
  const double jcN = jrow[(NS - 1) * 64];
  const double rho = s.rho, Wbar = s.Wbar, T = s.T;
  const double invWN = t.sp[(K - 1) * kSpStride];
  const double Wk = t.sp[((lane >= 1 && lane < K) ? lane - 1 : 0) * kSpStride + 1];
  const double fac = Wk / rho;
  const bool isT = lane == 0;
  const bool live = lane < NS;
  const double D = rho * cpmix;
  const double invD = 1.0 / D;
  const double fT = -Q * invD;
  const double cpN = w.cp[K - 1];
  // Chain rule to (T, Y). Constant pressure: C_i = P Y_i / (R T Wbar W_i) and m = P / (R T)
  // move with T (-C/T, -m/T) and C with Y through Wbar (-Wbar d_j C_i), and the species rows'
  // 1/rho moves too (+wdot/T, +wdot Wbar d_j). Constant volume: C_i = rho Y_i / W_i does not
  // depend on T, m = rho sum Y/W gives dm/dY_j = rho d_j, and rho is fixed.
  const double cpf = 1.0 - cvf;
  // column 0 (temperature)
  {
    double dWdT = JT_k - cpf * (sigma * s.invT + (s.m * s.invT) * Jm_k);
    double spv = fac * (dWdT + cpf * wdot_k * s.invT);
    double dQdT = SCpw + SHJT - cpf * (sigma * s.invT + (s.m * s.invT) * SHJm);  // lane 0: sigma = sum_k H_k sigma_k
    double dDdT = -cpf * rho * cpmix * s.invT + rho * dcpmix;
    double tv = (-dQdT - fT * dDdT) * invD;
    if (live) g.J[lane] = isT ? tv : spv;
  }
  const double c1 = jcN * rho * invWN;
  // per column j: base += d_j * ysig, species rows add d_j * wcor, D's derivative d_j * dcoef
  const double ysig = t.conv ? rho * (isT ? SHJm : Jm_k) : -Wbar * sigma;
  const double wcor = cpf * wdot_k * Wbar;
  const double dcoef = -cpf * rho * Wbar * cpmix;
#pragma unroll 2
  for (int j = 0; j < NS - 1; ++j) {
    const double invWj = t.sp[j * kSpStride];
    const double dj = invWj - invWN;
    const double base = jrow[j * 64] * rho * invWj - c1 + dj * ysig;
    double spv = fac * (base + wcor * dj);
    double dDdY = dcoef * dj + rho * (w.cp[j] - cpN);
    double tv = (-base - fT * dDdY) * invD;
    if (live) g.J[(j + 1) * NS + lane] = isT ? tv : spv;
  }
  wave_sync();
}

// ---------------------------------------------------------------------------
// Dense LU of the Newton matrix, one row per lane (SUNDIALS denseGETRF order of
// operations with LAPACK-position pivot ties), and the matching solve.
// ---------------------------------------------------------------------------
HPO_DI double pivot_key(double v, int pos) {
  unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(fabs(v)));
  b = (b & ~63ull) | static_cast<unsigned long long>(63 - pos);
  return __longlong_as_double(static_cast<long long>(b));
}
HPO_DI float pivot_key(float v, int pos) {
  unsigned b = static_cast<unsigned>(__float_as_int(fabsf(v)));
  b = (b & ~63u) | static_cast<unsigned>(63 - pos);
  return __int_as_float(static_cast<int>(b));
}
HPO_DI int key_pos(double k) { return 63 - static_cast<int>(__double_as_longlong(k) & 63); }
HPO_DI int key_pos(float k) { return 63 - (__float_as_int(k) & 63); }

// Row layout of the factors: after the elimination every row moves to the lane of its
// elimination step (one lane permutation per register), so both triangular solves broadcast
// from fixed lanes. Lane k then holds row k of the factors in pivot order: a[j < k] = L[k][j]
// (unit lower triangle), a[j > k] = U[k][j] / U[k][k] (upper triangle scaled to a unit
// diagonal), and rdiag = 1 / U[k][k]. The scaled upper triangle lets the back substitution
// read each x_k with one broadcast and no multiply.
template <int NS, typename F>
struct LuReg {
  F a[NS];
  F rdiag;    // lane k: 1 / U[k][k]
  int perm;   // lane k: the lane (original row) that pivoted at step k
};

HPO_DI float shfl(float v, int src) { return __shfl(v, src, 64); }

// Branch-free over the elimination steps: a zero or NaN pivot column marks the matrix
// singular and turns the remaining updates into no-ops instead of leaving the loop, so the
// unrolled elimination has a single exit. The lane index is re-hidden from the optimiser on
// entry so the per-step lane comparisons are formed here, not hoisted out of the caller's
// loops into dozens of long-lived masks.
template <int NS, typename F>
HPO_DI bool lu_factor(LuReg<NS, F>& L, int lane_in) {
  const int lane = opaque(lane_in);
  int pos = lane;
  bool done = false;
  bool singular = false;
  F rd = F(0);
  int perm = 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const bool cand = (!done) && (lane < NS);
    F key = cand ? pivot_key(L.a[k], pos) : F(0);
    F kmax = wmax(key);
    int ppos = key_pos(kmax);
    unsigned long long bal = __ballot(cand && pos == ppos);
    int p = bal ? (__ffsll(bal) - 1) : 0;
    F piv = rl(L.a[k], p);
    bool bad = (bal == 0ull) || (piv == F(0));
    singular = singular || bad;
    F mult = bad ? F(0) : F(1) / piv;
    if (!bad) {
      if (pos == k) pos = ppos;
      if (lane == p) {
        pos = k;
        done = true;
        rd = mult;
      }
    }
    perm = wli(perm, p, k, lane);
    F l = done ? F(0) : L.a[k] * mult;
    if (!done) L.a[k] = l;
#pragma unroll
    for (int j = k + 1; j < NS; ++j) {
      F pj = rl(L.a[j], p);
      L.a[j] = fma(-l, pj, L.a[j]);
    }
  }
  const bool live = lane < NS;
  const int src = live ? perm : lane;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    F v = live ? shfl(L.a[j], src) : F(0);
    L.a[j] = v;
  }
  rd = live ? shfl(rd, src) : F(0);
#pragma unroll
  for (int j = 1; j < NS; ++j) L.a[j] = (lane < j) ? L.a[j] * rd : L.a[j];
  L.rdiag = rd;
  L.perm = perm;
  return !singular;
}

// Solves (L U) x = P b for the factors of lu_factor; b and x are one entry per lane in the
// original row / column order. Each substitution step updates every lane without a lane mask:
// the entry finished at step k is first copied into the result with v_writelane, so the
// meaningless updates of already-finished lanes (and of the lane itself) are never read.
template <int NS, typename F>
HPO_DI F lu_solve(const LuReg<NS, F>& L, F b, int lane_in) {
  const int lane = opaque(lane_in);
  const bool live = lane < NS;
  b = shfl(b, live ? L.perm : lane);
#pragma unroll
  for (int k = 0; k < NS - 1; ++k) {
    F yk = rl(b, k);
    F c = (lane > k) ? L.a[k] : F(0);
    b = fma(-c, yk, b);
  }
  // U' x = D^-1 y with the unit-diagonal U' = D^-1 U
  b *= L.rdiag;
#pragma unroll
  for (int k = NS - 1; k >= 1; --k) {
    F xk = rl(b, k);
    F c = (lane < k) ? L.a[k] : F(0);
    b = fma(-c, xk, b);
  }
  return live ? b : F(0);
}

// ---------------------------------------------------------------------------
// FP32 Newton matrix: Gauss-Jordan inverse with the same partial pivoting as lu_factor
// (LAPACK's first-maximum choice among the rows not yet pivoted, so the pivot sequence is
// the LU's), one row per lane. A solve is then a matrix-vector product whose operand is
// broadcast through LDS: no serial substitution chain.
// ---------------------------------------------------------------------------
// Unscaled in-place elimination: at step k every other row (earlier pivot rows included)
// subtracts l = a[k] / pivot times the pivot row, the pivot row keeps its values, and column
// k takes the inverse's entries (-l; 1 in the pivot row). At the end each row is divided by
// its own pivot and moved to the lane of its step, so lane k holds row k of (P M)^-1 and
// M^-1 b = (P M)^-1 (P b).
//
// The pivot search is an unsigned maximum over 32-bit keys (|a| with the row position in the
// low six bits): integer maxima need no NaN canonicalisation and fuse with their DPP moves.
HPO_DI unsigned pivot_key_u(float v, int pos) {
  return (static_cast<unsigned>(__float_as_int(fabsf(v))) & ~63u) | static_cast<unsigned>(63 - pos);
}
template <int CTRL>
HPO_DI unsigned dpp_u(unsigned v) {
  return static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
}
HPO_DI unsigned wmax_u(unsigned v) {
  v = max(v, dpp_u<0xB1>(v));
  v = max(v, dpp_u<0x4E>(v));
  v = max(v, dpp_u<0x141>(v));
  v = max(v, dpp_u<0x140>(v));
  const unsigned a = static_cast<unsigned>(rli(static_cast<int>(v), 0));
  const unsigned b = static_cast<unsigned>(rli(static_cast<int>(v), 16));
  const unsigned c = static_cast<unsigned>(rli(static_cast<int>(v), 32));
  const unsigned d = static_cast<unsigned>(rli(static_cast<int>(v), 48));
  return max(max(a, b), max(c, d));
}

template <int NS>
HPO_DI bool gj_invert(LuReg<NS, float>& L, int lane_in) {
  const int lane = opaque(lane_in);
  int pos = lane;
  bool done = false;
  bool singular = false;
  float rd = 0.f;
  int perm = 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const bool cand = (!done) && (lane < NS);
    const unsigned key = cand ? pivot_key_u(L.a[k], pos) : 0u;
    const unsigned kmax = wmax_u(key);
    const int ppos = 63 - static_cast<int>(kmax & 63u);
    unsigned long long bal = __ballot(cand && pos == ppos);
    int p = bal ? (__ffsll(bal) - 1) : 0;
    float piv = rl(L.a[k], p);
    bool bad = (bal == 0ull) || !(fabsf(piv) > 0.f);
    singular = singular || bad;
    // hardware reciprocal (1 ulp): the FP32 inverse is refined against the FP64 matrix
    float mult = bad ? 0.f : __builtin_amdgcn_rcpf(piv);
    const bool isp = lane == p;
    if (!bad) {
      if (pos == k) pos = ppos;
      if (isp) {
        pos = k;
        done = true;
        rd = mult;
      }
    }
    perm = wli(perm, p, k, lane);
    const float l = isp ? 0.f : L.a[k] * mult;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      if (j == k) continue;
      float pj = rl(L.a[j], p);
      L.a[j] = fma(-l, pj, L.a[j]);
    }
    L.a[k] = isp ? 1.f : -l;
  }
  const bool live = lane < NS;
  const int src = live ? perm : lane;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    float v = L.a[j] * rd;
    L.a[j] = live ? shfl(v, src) : 0.f;
  }
  L.rdiag = 0.f;
  L.perm = perm;
  return !singular;
}

// FP32 solve of the BDF kernel's Newton system in either layout: the explicit inverse of
// gj_invert (buf: 64 floats of this wave's LDS) or the in-place factors of lu_factor.
template <int NS, bool INV>
HPO_DI float solve32(const LuReg<NS, float>& L, float b, int lane, float* buf);

// x = M^-1 b for the inverse of gj_invert; b and x are one entry per lane in the original
// order. buf: 64 floats of this wave's LDS.
template <int NS>
HPO_DI float gj_apply(const LuReg<NS, float>& L, float b, int lane_in, float* buf) {
  const int lane = opaque(lane_in);
  const bool live = lane < NS;
  const float bp = shfl(b, live ? L.perm : lane);
  wave_sync();
  buf[lane] = bp;
  wave_sync();
  float x0 = 0.f, x1 = 0.f;
#pragma unroll
  for (int j = 0; j + 1 < NS; j += 2) {
    x0 = fma(L.a[j], buf[j], x0);
    x1 = fma(L.a[j + 1], buf[j + 1], x1);
  }
  if (NS & 1) x0 = fma(L.a[NS - 1], buf[NS - 1], x0);
  return live ? x0 + x1 : 0.f;
}

template <int NS, bool INV>
HPO_DI float solve32(const LuReg<NS, float>& L, float b, int lane, float* buf) {
  if constexpr (INV)
    return gj_apply<NS>(L, b, lane, buf);
  else
    return lu_solve<NS>(L, b, lane);
}

}  // namespace dev
}  // namespace hpo
