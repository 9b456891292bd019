// Chemistry mechanism: host tables read from a Cantera YAML file, and the packed
// device tables the HIP kernels stage into LDS.
//
// Replaces the reference's pyJac code generation (pyJac_mechanism_generator.sh:3-4,
// src/Mechanism/CPU/*.c, src/Mechanism/GPU/*.cu): the same physics is evaluated from
// tables at run time, so any mechanism file in this format can be loaded without
// regenerating code. Conventions follow the generated code:
//   R = 8314.4621 J/(kmol K) (chem_utils.c:42), A in SI-kmol, theta = Ea/R,
//   NASA7 low range for T <= T_mid (chem_utils.c:197), third-body [M] = base*P/(RT)
//   + sum (eff-base) C (rxn_rates_pres_mod.c:20-23), Troe falloff
//   (rxn_rates_pres_mod.c:65-73), the last species eliminated (mechanism.h:5).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hpo {

constexpr double kRGas = 8314.4621;  // J/(kmol K)
constexpr double kPAtm = 101325.0;   // Pa
constexpr int kMaxSideSpecies = 3;   // species per reaction side on the device path

enum RxnKind : int { kElementary = 0, kThreeBody = 1, kLindemann = 2, kTroe = 3 };

// One pressure level of a pressure-dependent Arrhenius rate (PLOG, Cantera's PlogRate).
struct PlogLevel {
  double lnP = 0, lnA = 0, b = 0, theta = 0, A = 0;   // ln of the pressure in Pa; SI-kmol Arrhenius
};

struct Reaction {
  std::string equation;
  int kind = kElementary;
  bool reversible = true;
  std::vector<std::pair<int, double>> reactants, products;  // (species, stoich)
  double lnA = 0, b = 0, theta = 0, A = 0;                 // high-pressure / plain rate
  double lnA0 = 0, b0 = 0, theta0 = 0, A0 = 0;             // low-pressure limit
  std::vector<std::pair<int, double>> eff;                 // (species, efficiency - base)
  double eff_base = 1.0;
  double troe[4] = {0, 0, 0, 0};                           // a, T3, T1, T2
  bool troe_has_t2 = false;
  // PLOG (kind stays kElementary): ln k interpolated linearly in ln P between the levels that
  // bracket the cell's pressure, the end level's rate outside them; levels ascending in P.
  std::vector<PlogLevel> plog;
};

struct Mechanism {
  std::string path;
  std::vector<std::string> species;
  std::vector<double> W, tmid;        // [K]
  std::vector<double> nasa;           // [K][14]: low a0..a6, high a0..a6
  std::vector<Reaction> reactions;
  int K() const { return static_cast<int>(species.size()); }
  int R() const { return static_cast<int>(reactions.size()); }
};

Mechanism load_mechanism_yaml(const std::string& path);
Mechanism load_mechanism_xml(const std::string& path);   // Cantera CTML
Mechanism load_mechanism_cti(const std::string& path);   // Cantera CTI
Mechanism load_mechanism(const std::string& path);       // by extension: .xml/.ctml, .cti or YAML

constexpr double kGasConstant = 8314.4621;   // J/(kmol K), the value the kernels use

// Mass-specific enthalpy h(T, Y) = sum_k Y_k H_k(T) / W_k in J/kg, H_k from the NASA7
// polynomials; Y holds all K mass fractions.
double enthalpy_mass(const Mechanism& m, double T, const double* Y);

// ---------------------------------------------------------------------------
// Device tables. One contiguous blob; the kernel copies the "LDS part" into
// shared memory once per workgroup and reads the "global part" (Jacobian
// assembly lists) through the cache.
// ---------------------------------------------------------------------------
constexpr uint8_t kDummySpecies = 63;  // unused reaction slots point here (concentration 1, g 0)

struct alignas(16) RxnHdr {     // 16 bytes
  uint8_t sp[6];                // reactants 0..2, products 3..5 (unused: kDummySpecies with nu = 0)
  uint8_t nu[6];                // integer stoichiometric coefficients
  uint8_t kind;                 // RxnKind
  uint8_t flags;                // bit0 reversible, bit1 constant kf (=A), bit2 Troe T2, bit3 constant k0
  uint16_t pidx;                // index into pressure-dependent tables (pressure-dependent reactions)
};

// The same reaction as molecules for the rate evaluation: each side lists up to three
// molecules (a species with coefficient 2 appears twice), unused slots hold kDummySpecies,
// so the concentration products and Gibbs sums need no exponents or selects.
struct alignas(8) RxnMol {      // 8 bytes
  uint8_t m[6];                 // reactant molecules 0..2, product molecules 3..5
  int8_t dnu;                   // product molecules - reactant molecules
  uint8_t pad;
};

struct alignas(8) PdepRec {     // per pressure-dependent reaction
  double eff_base;
  int32_t eff_begin, eff_end;   // range in eff tables
  double lnA0, b0, theta0;      // low-pressure limit (falloff)
  double troe_a, troe_iT3, troe_iT1, troe_T2;   // a, 1/T3, 1/T1 (0 -> inf: that term vanishes), T2
};

// Layout of the blob, offsets in bytes from its start. Sizes in elements.
struct DevMechLayout {
  int32_t K, R, n_pdep, n_eff;
  int32_t n_gather_lane;        // entries per lane in the balanced species gather
  int32_t n_partials;           // partial-sum slots of the gather
  int32_t n_slots;              // d(phi)/dC slots (Jacobian)
  int32_t n_jterm_lane;         // Jacobian assembly terms per lane
  int32_t n_coef;               // distinct stoichiometric coefficient values
  int32_t n_eff_pad;            // efficiency entries per pressure-dependent reaction in the padded table
  int32_t mol_ok;               // 1: every reaction side has at most 3 molecules (RxnMol table valid)
  int32_t lds_bytes;            // bytes [0, lds_bytes) are staged to LDS
  // LDS part
  int32_t off_species;          // double[K][16]: invW, W, tmid, (pad), nasa[14] -> see kSp*
  int32_t off_hdr;              // RxnHdr[R]
  int32_t off_mol;              // RxnMol[R]
  int32_t off_arr;              // double[R][3]: lnA (or A), b, theta
  int32_t off_pdep;             // PdepRec[n_pdep]
  int32_t off_eff_val;          // double[n_eff]
  int32_t off_eff_sp;           // uint8[n_eff] (padded to 8)
  int32_t off_effp_val;         // double[n_pdep][n_eff_pad]: efficiencies, padded with 0.0
  int32_t off_effp_sp;          // uint8[n_pdep][n_eff_pad]: their species, padded with kDummySpecies
  int32_t off_coef;             // double[16] distinct coefficient values
  int32_t off_gather;           // uint32[n_gather_lane][64]  r | nu<<12 | flush<<16 | part<<17
                                //   (nu: net stoichiometric coefficient, 4-bit two's complement)
  int32_t off_part_ptr;         // uint16[K+1] partial-slot range per species
  int32_t off_exptab;           // double[32]: 2^(j/32), the table of the kernels' exponential
  // global part
  int32_t off_slot_base;        // uint16[R+1]: first d(phi)/dC slot of each reaction
  int32_t off_jterm;            // uint64[n_jterm_lane][64]  cell | src<<16 | coef<<32 | last<<48
  int32_t total_bytes;
};

constexpr int kSpStride = 17;   // doubles per species record: invW, W, tmid, nasa[14]

struct DevMechBlob {
  DevMechLayout lay{};
  std::vector<uint8_t> bytes;
};

// Builds the packed tables. Throws std::runtime_error when the mechanism does not fit
// the device path (more than 64 state variables, >3 species on a reaction side,
// non-integer stoichiometry).
DevMechBlob build_device_tables(const Mechanism& m);

// ---------------------------------------------------------------------------
// Block-per-cell tables: one workgroup integrates one cell, thread t owning state component t,
// so the species count is bounded by the workgroup (and 255 by the byte species indices), not
// by 64 lanes. Species production rates and Jacobian rows are gathered per species from a
// compressed row list (owner computes, fixed order), the temperature row per column from a
// list of the Jacobian slots of each species.
// ---------------------------------------------------------------------------
constexpr int kBlkSide = 4;     // species per reaction side

struct alignas(8) BlkRxn {      // 24 bytes
  uint8_t sp[8];                // reactants 0..3, products 4..7; unused slots: the dummy species K
  uint8_t nu[8];                // integer stoichiometric coefficients (0 in unused slots)
  uint8_t kind;                 // RxnKind
  uint8_t flags;                // as RxnHdr::flags
  uint16_t pidx;                // pressure-dependent record
  uint16_t slot;                // first d(phi)/dC slot of the reaction
  uint16_t nslot;               // its slots: reactants, products (reversible), efficiencies
};
static_assert(sizeof(BlkRxn) == 24, "BlkRxn layout");

struct BlkLayout {
  int32_t K, R, n_pdep, n_eff, n_slots, n_csr, n_col, n_coef, n_plog, n_plog_lv, n_lump;
  // Bytes [0, lds_bytes) are staged to LDS: a prefix of the tables below, in this order, as many
  // as fit beside the kernel's work area (always through the Arrhenius table, which PLOG rewrites
  // per cell); the others are read from global memory through the cache.
  int32_t lds_bytes;
  int32_t off_exptab;           // double[32]: 2^(j/32)
  int32_t off_arr;              // double[R][3]: lnA (or A), b, theta; PLOG reactions: per cell
  int32_t off_species;          // double[K][kSpStride]
  int32_t off_rxn;              // BlkRxn[R]
  int32_t off_pdep;             // PdepRec[n_pdep]
  int32_t off_eff_val;          // double[n_eff]
  int32_t off_eff_sp;           // uint8[n_eff]
  int32_t off_csr_ptr;          // uint16[K+1]: reactions of species k
  int32_t off_coef;             // double[n_coef]: the distinct net stoichiometric coefficients
  int32_t off_csr;              // uint32[n_csr]: r | coefficient index << 16, ascending r
  // global only
  int32_t off_slot_sp;          // uint8[n_slots]: species of each d(phi)/dC slot
  int32_t off_col_ptr;          // uint32[K+1]: Jacobian slots whose species is k
  int32_t off_col;              // uint32[n_col]: slot | reaction << 16
  int32_t off_plog_rxn;         // uint16[n_plog]: device reaction of each PLOG rate
  int32_t off_plog_ptr;         // uint16[n_plog+1]: its pressure levels
  int32_t off_plog_lv;          // double[n_plog_lv][4]: ln P (Pa), lnA, b, theta, ascending in P
  int32_t off_lump_ptr;         // uint16[R+1]: products of lumped reactions (flag 16)
  int32_t off_lump;             // uint32[n_lump]: species | coefficient index << 16
  int32_t total_bytes;
};

constexpr int kBlkLdsMax = 160 * 1024;   // LDS per CU (gfx950)
constexpr uint8_t kBlkLumped = 16;       // BlkRxn::flags: products in the lump list, not the slots

struct DevBlkBlob {
  BlkLayout lay{};
  std::vector<uint8_t> bytes;
};

// work_lds_bytes: the kernel's LDS work area beside the tables, which bounds the staged prefix.
// Throws std::runtime_error when the mechanism does not fit: more than 255 species, more than 4
// reactant species or 4 products of a reversible reaction, non-integer reactant or reversible
// product stoichiometry, more than 65535 reactions, slots or coefficients, or an LDS work area
// with no room for the Arrhenius table. Irreversible reactions may have any number of products
// with any positive coefficient ("lumped": products enter the production rates and the reaction
// enthalpy only).
DevBlkBlob build_block_tables(const Mechanism& m, int work_lds_bytes);

}  // namespace hpo
