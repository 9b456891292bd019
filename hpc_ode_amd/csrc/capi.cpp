// C ABI implementation (include/hpc_ode.h): handles, device memory, launches.
#include "../../include/hpc_ode.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_internal.hpp"

namespace hpo {
namespace capi {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int map_exceptions() {
  try {
    throw;
  } catch (const HipError& e) {
    return fail(e.e == hipErrorOutOfMemory ? HPO_E_NOMEM : HPO_E_DEVICE, e.what());
  } catch (const std::invalid_argument& e) {
    return fail(HPO_E_INVALID, e.what());
  } catch (const std::exception& e) {
    return fail(HPO_E_PARSE, e.what());
  } catch (...) {
    return fail(HPO_E_DEVICE, "unknown error");
  }
}

}  // namespace capi
}  // namespace hpo

using namespace hpo::capi;

hpo_mech::~hpo_mech() {
  if (device >= 0) {
    (void)hipSetDevice(device);
    if (d_blob) (void)hipFree(d_blob);
    if (d_bblob) (void)hipFree(d_bblob);
    if (d_scratch) (void)hipFree(d_scratch);
    if (d_queue) (void)hipFree(d_queue);
    if (d_backup) (void)hipFree(d_backup);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_last) (void)hipEventDestroy(ev_last);
  }
}

void hpo_mech::ensure_device() {
  if (d_queue) return;
  HPO_HIP(hipGetDevice(&device));
  hipDeviceProp_t prop;
  HPO_HIP(hipGetDeviceProperties(&prop, device));
  n_cu = prop.multiProcessorCount;
  if (has_wave) {
    d_blob = dalloc<uint8_t>(blob.bytes.size());
    HPO_HIP(hipMemcpy(d_blob, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice));
    scratch_stride = hpo::wave_scratch_doubles_host(blob.lay);
  }
  if (has_block) {
    d_bblob = dalloc<uint8_t>(bblob.bytes.size());
    HPO_HIP(hipMemcpy(d_bblob, bblob.bytes.data(), bblob.bytes.size(), hipMemcpyHostToDevice));
  }
  d_queue = dalloc<uint32_t>(4);
  HPO_HIP(hipEventCreate(&ev0));
  HPO_HIP(hipEventCreate(&ev1));
  HPO_HIP(hipEventCreateWithFlags(&ev_last, hipEventDisableTiming));
}

void hpo_mech::ensure_scratch_doubles(int64_t n) {
  if (n <= scratch_cap) return;
  // the previous launch on this scratch must be done before it is freed
  if (launched) HPO_HIP(hipEventSynchronize(ev_last));
  if (d_scratch) HPO_HIP(hipFree(d_scratch));
  d_scratch = nullptr;
  d_scratch = dalloc<double>(static_cast<size_t>(n));
  scratch_cap = n;
}

void hpo_mech::ensure_scratch(int64_t slots) { ensure_scratch_doubles(slots * scratch_stride); }

void hpo_mech::ensure_backup_doubles(int64_t n) {
  if (n <= backup_cap) return;
  if (launched) HPO_HIP(hipEventSynchronize(ev_last));
  if (d_backup) HPO_HIP(hipFree(d_backup));
  d_backup = nullptr;
  d_backup = dalloc<double>(static_cast<size_t>(n));
  backup_cap = n;
}

struct hpo_batch {
  const hpo_mech* mech = nullptr;
  int64_t capacity = 0;
  int64_t ld = 0;
  int device = -1;
  double *d_y = nullptr, *d_P = nullptr, *d_dt = nullptr;
  int32_t *d_status = nullptr, *d_stats = nullptr, *d_order = nullptr;
  int64_t tuned_count = -1;   // cells the order in d_order was tuned for (-1: none)
  ~hpo_batch() {
    if (device >= 0) {
      (void)hipSetDevice(device);
      for (void* p : {static_cast<void*>(d_y), static_cast<void*>(d_P), static_cast<void*>(d_dt),
                      static_cast<void*>(d_status), static_cast<void*>(d_stats), static_cast<void*>(d_order)})
        if (p) (void)hipFree(p);
    }
  }
};

namespace hpo {
namespace capi {

void check_opts(const hpo_opts& o) {
  if (!(o.rtol >= 0.0) || !(o.atol >= 0.0)) throw std::invalid_argument("tolerances must be >= 0");
  if (o.lu_precision != 0 && o.lu_precision != 1) throw std::invalid_argument("lu_precision must be 0 or 1");
  if (o.sort_cells < 0 || o.sort_cells > 2) throw std::invalid_argument("sort_cells must be 0, 1 or 2");
  if (o.refine_sweeps < 0 || o.refine_sweeps > 8) throw std::invalid_argument("refine_sweeps must be in 0..8");
  if (o.waves_per_block < 0 || o.waves_per_block > 4) throw std::invalid_argument("waves_per_block must be in 0..4");
  if (o.blocks_per_cu < 0) throw std::invalid_argument("blocks_per_cu must be >= 0");
  if (o.lu_layout != 0 && o.lu_layout != 1) throw std::invalid_argument("lu_layout must be 0 or 1");
  if (o.cell_layout < 0 || o.cell_layout > 2) throw std::invalid_argument("cell_layout must be 0, 1 or 2");
  if (o.reactor < 0 || o.reactor > 1) throw std::invalid_argument("reactor must be 0 (constant pressure) or 1 (constant volume)");
  if (o.fp64_fallback != 0 && o.fp64_fallback != 1) throw std::invalid_argument("fp64_fallback must be 0 or 1");
}

bool use_block_layout(const hpo_mech* m, const hpo_opts& o) {
  const int ns = m->host.K();
  const bool wave_ok = m->has_wave && hpo::kernel_supports(ns);
  const bool block_ok = m->has_block && hpo::block_kernel_supports(ns);
  if (o.cell_layout == 1) {
    if (!wave_ok) throw std::invalid_argument("no wave-per-cell kernel for a state of " + std::to_string(ns) + " variables");
    return false;
  }
  if (o.cell_layout == 2) {
    if (!block_ok) throw std::invalid_argument("no block-per-cell kernel for a state of " + std::to_string(ns) + " variables");
    return true;
  }
  if (wave_ok) return false;
  if (block_ok) return true;
  throw std::invalid_argument("no kernel instantiated for a state of " + std::to_string(ns) + " variables");
}

// Whether two arrays of n doubles share memory.
static bool rows_overlap(const double* a, const double* b, int64_t n) {
  const uintptr_t ua = reinterpret_cast<uintptr_t>(a), ub = reinterpret_cast<uintptr_t>(b);
  const uintptr_t bytes = static_cast<uintptr_t>(n) * sizeof(double);
  return ua < ub + bytes && ub < ua + bytes;
}

// The FP64 launch that follows an FP32 one with the fallback armed (queue word 3 set by
// launch_reset_counter): count_failed, then the same arguments with the input rows at
// d_rows_in (the FP32 launch's own input, or its copy when it ran in place), on a grid that fits
// the FP64 kernel's occupancy, the share of the device and the scratch the FP32 launch used (the
// FP64 kernel never holds more waves per CU than the FP32 one, so the grid is rarely cut by it).
static void launch_fp64_fallback(const hpo_mech* m, bool block_layout, hpo::BdfArgs A, const double* d_rows_in,
                                 int wpb, int block, size_t lds, int share_pct, int64_t scratch_doubles,
                                 hipStream_t st) {
  const int ns = m->host.K();
  HPO_HIP(hpo::launch_count_failed(A.status, A.ncells, A.queue, st));
  const bool prof = A.prof != nullptr;
  const void* fn = block_layout ? hpo::bdf_block_kernel_fn(ns, 0, prof) : hpo::bdf_kernel_fn(ns, 0, prof, A.reactor);
  int occ = 0;
  HPO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, block, lds));
  if (occ < 1) throw std::invalid_argument("the FP64 kernel does not fit on a CU with this configuration");
  int64_t grid = static_cast<int64_t>(m->n_cu) * occ;
  if (share_pct < 100) grid = (grid * std::max(share_pct, 1) + 99) / 100;
  grid = std::min<int64_t>(grid, (A.ncells + wpb - 1) / wpb);
  grid = std::min<int64_t>(grid, scratch_doubles / (static_cast<int64_t>(wpb) * A.scratch_stride));
  grid = std::max<int64_t>(grid, 1);
  A.order = nullptr;
  A.y_in = d_rows_in;
  if (block_layout)
    HPO_HIP(hpo::launch_bdf_block(ns, A, 0, static_cast<int>(grid), lds, st));
  else
    HPO_HIP(hpo::launch_bdf(ns, A, 0, static_cast<int>(grid), block, lds, st));
}

// Launches the persistent BDF kernel on arrays already resident on the device.
void launch_integrate(hpo_mech* m, int64_t ncells, const double* d_yin, double* d_yout, int64_t ld,
                      const double* d_P, const double* d_dt, int32_t* d_status, int32_t* d_stats,
                      const int32_t* d_order, int64_t* d_prof, const hpo_opts& o, hipStream_t st,
                      hpo_timing* tm, const ExecRes* res) {
  check_opts(o);
  const int ns = m->host.K();
  const bool block_layout = use_block_layout(m, o);
  if (o.reactor == 1 && m->has_block && m->bblob.lay.n_plog > 0)
    throw std::invalid_argument("PLOG rates are supported at constant pressure only");
  if (ncells < 0 || ncells > 0x7fffffffLL) throw std::invalid_argument("ncells out of range");
  if (ld < ns) throw std::invalid_argument("leading dimension smaller than the state");
  m->ensure_device();
  int wpb, block;
  size_t lds;
  const void* fn;
  const int variant = hpo::bdf_variant(o.lu_precision, o.lu_layout);
  if (block_layout) {
    wpb = 1;   // one cell per 256-thread workgroup
    block = 256;
    lds = static_cast<size_t>(m->bblob.lay.lds_bytes) + hpo::blk_work_lds_bytes(m->bblob.lay);
    fn = hpo::bdf_block_kernel_fn(ns, o.lu_precision, d_prof != nullptr);
  } else {
    wpb = o.waves_per_block > 0 ? o.waves_per_block : 4;
    block = 64 * wpb;
    lds = static_cast<size_t>(m->blob.lay.lds_bytes) + static_cast<size_t>(wpb) * hpo::bdf_wave_lds_doubles(m->blob.lay) * 8;
    fn = hpo::bdf_kernel_fn(ns, variant, d_prof != nullptr, o.reactor);
  }
  int occ = 0;
  HPO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, block, lds));
  if (occ < 1) throw std::invalid_argument("kernel does not fit on a CU with this configuration");
  int bpc = o.blocks_per_cu > 0 ? std::min(o.blocks_per_cu, occ) : occ;
  int64_t grid = static_cast<int64_t>(m->n_cu) * bpc;
  if (res && res->share_pct < 100) grid = (grid * std::max(res->share_pct, 1) + 99) / 100;
  grid = std::min<int64_t>(grid, (ncells + wpb - 1) / wpb);
  grid = std::max<int64_t>(grid, 1);
  // scratch: per wave slot (wave layout) or per workgroup (block layout)
  const int64_t stride = block_layout ? hpo::blk_scratch_doubles_host(m->bblob.lay) : m->scratch_stride;
  const int64_t need = grid * wpb * stride;
  ExecRes own;
  std::unique_lock<std::mutex> own_lock;
  bool own_ctx = false, capturing = false;
  if (!res) {
    // the handle's own context: one launch at a time on the device (see hpo_mech::ev_last)
    own_lock = std::unique_lock<std::mutex>(m->launch_mu);
    own_ctx = true;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HPO_HIP(hipStreamIsCapturing(st, &cs));
    capturing = cs != hipStreamCaptureStatusNone;
    m->ensure_scratch_doubles(need);
    own.d_queue = m->d_queue;
    own.d_scratch = m->d_scratch;
    own.scratch_doubles = m->scratch_cap;
    res = &own;
    // a captured launch is not ordered against eager ones here: its replays must not overlap
    // other launches on this handle (the header says so)
    if (!capturing && m->launched && m->last_stream != st) HPO_HIP(hipStreamWaitEvent(st, m->ev_last, 0));
  }
  if (res->scratch_doubles < need) throw std::invalid_argument("launch context scratch too small for this launch");

  hpo::BdfArgs A{};
  A.mech = block_layout ? m->d_bblob : m->d_blob;
  A.lay = m->blob.lay;
  A.blay = m->bblob.lay;
  A.y_in = d_yin;
  A.y_out = d_yout;
  A.P = d_P;
  A.dt = d_dt;
  A.status = d_status;
  A.stats = d_stats;
  A.order = d_order;
  A.queue = res->d_queue;
  A.scratch = res->d_scratch;
  A.scratch_stride = stride;
  A.ncells = ncells;
  A.ld = ld;
  A.rtol = o.rtol;
  A.atol = o.atol;
  A.maxsteps = o.maxsteps;
  A.refine_sweeps = o.refine_sweeps;
  A.prof = reinterpret_cast<int64_t*>(d_prof);
  A.reactor = o.reactor;
  const bool fallback = o.lu_precision == 1 && o.fp64_fallback && ncells > 0;
  // The FP64 fallback restarts failed cells from their input rows: an in-place launch keeps a
  // copy of them (the FP32 kernel writes every output row).
  const double* rows_in = d_yin;
  const int64_t rows = ncells * ld;
  if (fallback && rows_overlap(d_yin, d_yout, rows)) {
    double* backup = nullptr;
    if (own_ctx) {
      m->ensure_backup_doubles(rows);
      backup = m->d_backup;
    } else {
      if (!res->d_backup || res->backup_doubles < rows)
        throw std::invalid_argument("launch context has no room for the input rows of the FP64 fallback");
      backup = res->d_backup;
    }
    rows_in = backup;
  }

  HPO_HIP(hpo::launch_reset_counter(res->d_queue, fallback, st));
  if (tm) HPO_HIP(hipEventRecord(m->ev0, st));
  if (rows_in != d_yin) HPO_HIP(hipMemcpyAsync(const_cast<double*>(rows_in), d_yin, rows * sizeof(double),
                                                hipMemcpyDeviceToDevice, st));
  if (ncells > 0) {
    if (block_layout)
      HPO_HIP(hpo::launch_bdf_block(ns, A, o.lu_precision, static_cast<int>(grid), lds, st));
    else
      HPO_HIP(hpo::launch_bdf(ns, A, variant, static_cast<int>(grid), block, lds, st));
    if (fallback)
      launch_fp64_fallback(m, block_layout, A, rows_in, wpb, block, lds, res->share_pct, res->scratch_doubles, st);
  }
  if (tm) HPO_HIP(hipEventRecord(m->ev1, st));
  if (own_ctx && !capturing) {
    HPO_HIP(hipEventRecord(m->ev_last, st));
    m->last_stream = st;
    m->launched = true;
  }
  if (tm) {
    HPO_HIP(hipEventSynchronize(m->ev1));
    float ms = 0.f;
    HPO_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    tm->kernel_ms = ms;
    tm->ncells = ncells;
    tm->grid = static_cast<int32_t>(grid);
    tm->block = block;
    tm->lds_bytes = static_cast<int32_t>(lds);
    tm->fallback_cells = 0;
    if (fallback) {
      uint32_t q[2] = {0, 0};
      HPO_HIP(hipMemcpyAsync(q, res->d_queue, sizeof(q), hipMemcpyDeviceToHost, st));
      HPO_HIP(hipStreamSynchronize(st));
      tm->fallback_cells = static_cast<int32_t>(q[1]);
    }
  }
}

int64_t max_scratch_doubles(hpo_mech* m) {
  m->ensure_device();
  const int ns = m->host.K();
  int64_t most = 0;
  if (m->has_wave && hpo::kernel_supports(ns)) {
    // one-wave workgroups pack the most waves onto a CU; the launch grid never exceeds
    // n_CU x (workgroups per CU) x (waves per workgroup)
    for (int wpb = 1; wpb <= 4; ++wpb)
      for (int variant = 0; variant < 3; ++variant) {
        const size_t lds = static_cast<size_t>(m->blob.lay.lds_bytes) +
                           static_cast<size_t>(wpb) * hpo::bdf_wave_lds_doubles(m->blob.lay) * 8;
        int occ = 0;
        HPO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, hpo::bdf_kernel_fn(ns, variant), 64 * wpb, lds));
        most = std::max<int64_t>(most, static_cast<int64_t>(m->n_cu) * occ * wpb * m->scratch_stride);
      }
  }
  if (m->has_block && hpo::block_kernel_supports(ns)) {
    const size_t lds = static_cast<size_t>(m->bblob.lay.lds_bytes) + hpo::blk_work_lds_bytes(m->bblob.lay);
    for (int lu32 = 0; lu32 < 2; ++lu32) {
      int occ = 0;
      HPO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, hpo::bdf_block_kernel_fn(ns, lu32, 0), 256, lds));
      most = std::max<int64_t>(most, static_cast<int64_t>(m->n_cu) * occ * hpo::blk_scratch_doubles_host(m->bblob.lay));
    }
  }
  return std::max<int64_t>(most, 1);
}

}  // namespace capi
}  // namespace hpo

extern "C" {

void hpo_opts_default(hpo_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->rtol = 1.0e-6;
  o->atol = 1.0e-10;
  o->maxsteps = 10000;
  o->lu_precision = 0;
  o->refine_sweeps = 1;
  o->fp64_fallback = 1;
}

const char* hpo_last_error(void) { return g_err.c_str(); }

static int mech_load(const char* path, hpo_mech** out, bool by_extension) {
  if (!path || !out) return fail(HPO_E_INVALID, "null argument");
  *out = nullptr;
  try {
    auto m = std::make_unique<hpo_mech>();
    m->host = by_extension ? hpo::load_mechanism(path) : hpo::load_mechanism_yaml(path);
    // wave-per-cell tables (at most 63 species) and block-per-cell tables (at most 255); the
    // mechanism loads when either fits, and the launch picks the layout (hpo_opts::cell_layout)
    std::string wave_err, block_err;
    try {
      m->blob = hpo::build_device_tables(m->host);
      m->has_wave = true;
    } catch (const std::exception& e) {
      wave_err = e.what();
    }
    try {
      m->bblob = hpo::build_block_tables(m->host, hpo::blk_work_lds_bytes_for(m->host.K(), m->host.R()));
      m->has_block = true;
    } catch (const std::exception& e) {
      block_err = e.what();
    }
    if (!m->has_wave && !m->has_block) throw std::runtime_error(wave_err + "; " + block_err);
    *out = m.release();
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

int hpo_mech_load_yaml(const char* path, hpo_mech** out) { return mech_load(path, out, false); }

int hpo_mech_load(const char* path, hpo_mech** out) { return mech_load(path, out, true); }

void hpo_mech_destroy(hpo_mech* m) { delete m; }

int hpo_mech_info(const hpo_mech* m, int32_t* n_species, int32_t* n_state, int32_t* n_reactions) {
  if (!m) return fail(HPO_E_INVALID, "null mechanism");
  if (n_species) *n_species = m->host.K();
  if (n_state) *n_state = m->host.K();
  if (n_reactions) *n_reactions = m->host.R();
  return HPO_OK;
}

int hpo_mech_species_name(const hpo_mech* m, int32_t k, char* buf, int32_t buflen) {
  if (!m || !buf || buflen <= 0) return fail(HPO_E_INVALID, "bad argument");
  if (k < 0 || k >= m->host.K()) return fail(HPO_E_INVALID, "species index out of range");
  std::snprintf(buf, static_cast<size_t>(buflen), "%s", m->host.species[k].c_str());
  return HPO_OK;
}

int hpo_mech_molecular_weights(const hpo_mech* m, double* W) {
  if (!m || !W) return fail(HPO_E_INVALID, "bad argument");
  std::copy(m->host.W.begin(), m->host.W.end(), W);
  return HPO_OK;
}

int hpo_mech_enthalpy_mass(const hpo_mech* m, int64_t n, const double* T, const double* Y, int64_t ldY, double* h) {
  if (!m || n < 0 || (n > 0 && (!T || !Y || !h))) return fail(HPO_E_INVALID, "bad argument");
  if (ldY < m->host.K()) return fail(HPO_E_INVALID, "ldY smaller than the number of species");
  for (int64_t i = 0; i < n; ++i) h[i] = hpo::enthalpy_mass(m->host, T[i], Y + i * ldY);
  return HPO_OK;
}

int hpo_mech_tables(const hpo_mech* m, double* nasa, double* tmid, double* arrh, int32_t* kind, int32_t* rev) {
  if (!m) return fail(HPO_E_INVALID, "null mechanism");
  const int K = m->host.K(), R = m->host.R();
  if (nasa) std::copy(m->host.nasa.begin(), m->host.nasa.end(), nasa);
  if (tmid) std::copy(m->host.tmid.begin(), m->host.tmid.end(), tmid);
  for (int r = 0; r < R; ++r) {
    const hpo::Reaction& x = m->host.reactions[r];
    if (arrh) {
      double* a = arrh + 8 * r;
      a[0] = x.lnA;
      a[1] = x.b;
      a[2] = x.theta;
      a[3] = x.A;
      a[4] = x.kind >= hpo::kLindemann ? x.lnA0 : 0.0;
      a[5] = x.kind >= hpo::kLindemann ? x.b0 : 0.0;
      a[6] = x.kind >= hpo::kLindemann ? x.theta0 : 0.0;
      a[7] = x.kind >= hpo::kLindemann ? x.A0 : 0.0;
    }
    if (kind) kind[r] = x.kind;
    if (rev) rev[r] = x.reversible ? 1 : 0;
  }
  (void)K;
  return HPO_OK;
}

int hpo_batch_create(const hpo_mech* m, int64_t capacity, hpo_batch** out) {
  if (!m || !out || capacity <= 0) return fail(HPO_E_INVALID, "bad argument");
  *out = nullptr;
  try {
    auto b = std::make_unique<hpo_batch>();
    b->mech = m;
    b->capacity = capacity;
    b->ld = m->host.K();
    HPO_HIP(hipGetDevice(&b->device));
    b->d_y = dalloc<double>(static_cast<size_t>(capacity * b->ld));
    b->d_P = dalloc<double>(capacity);
    b->d_dt = dalloc<double>(capacity);
    b->d_status = dalloc<int32_t>(capacity);
    b->d_stats = dalloc<int32_t>(static_cast<size_t>(capacity * HPO_NSTATS));
    *out = b.release();
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

void hpo_batch_destroy(hpo_batch* b) { delete b; }

int hpo_batch_put(hpo_batch* b, int64_t first, int64_t count, const double* y, int64_t ldy, const double* P,
                  double P_scalar, const double* dt, double dt_scalar) {
  if (!b || !y) return fail(HPO_E_INVALID, "null argument");
  if (first < 0 || count < 0 || first + count > b->capacity) return fail(HPO_E_INVALID, "range exceeds capacity");
  if (ldy < b->ld) return fail(HPO_E_INVALID, "ldy smaller than the state length");
  if (count == 0) return HPO_OK;
  try {
    HPO_HIP(hipSetDevice(b->device));
    HPO_HIP(hipMemcpy2D(b->d_y + first * b->ld, b->ld * sizeof(double), y, ldy * sizeof(double),
                        b->ld * sizeof(double), count, hipMemcpyHostToDevice));
    std::vector<double> tmp;
    if (P) {
      HPO_HIP(hipMemcpy(b->d_P + first, P, count * sizeof(double), hipMemcpyHostToDevice));
    } else {
      tmp.assign(count, P_scalar);
      HPO_HIP(hipMemcpy(b->d_P + first, tmp.data(), count * sizeof(double), hipMemcpyHostToDevice));
    }
    if (dt) {
      HPO_HIP(hipMemcpy(b->d_dt + first, dt, count * sizeof(double), hipMemcpyHostToDevice));
    } else {
      tmp.assign(count, dt_scalar);
      HPO_HIP(hipMemcpy(b->d_dt + first, tmp.data(), count * sizeof(double), hipMemcpyHostToDevice));
    }
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

int hpo_batch_integrate(hpo_batch* b, int64_t count, const hpo_opts* o, void* stream, hpo_timing* tm) {
  if (!b) return fail(HPO_E_INVALID, "null batch");
  if (count < 0 || count > b->capacity) return fail(HPO_E_INVALID, "count exceeds capacity");
  hpo_opts def;
  hpo_opts_default(&def);
  const hpo_opts& opt = o ? *o : def;
  try {
    HPO_HIP(hipSetDevice(b->device));
    hpo_mech* m = const_cast<hpo_mech*>(b->mech);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int32_t* order = nullptr;
    if (opt.sort_cells == 2) {
      if (b->tuned_count != count) throw std::invalid_argument("sort_cells = 2 needs hpo_batch_autotune for this count");
      order = b->d_order;
    } else if (opt.sort_cells == 1 && count > 1) {
      // hottest cells first: they take the most steps, so the queue tail holds short cells
      std::vector<double> T(count);
      HPO_HIP(hipMemcpy2D(T.data(), sizeof(double), b->d_y, b->ld * sizeof(double), sizeof(double), count,
                          hipMemcpyDeviceToHost));
      std::vector<int32_t> idx(count);
      std::iota(idx.begin(), idx.end(), 0);
      std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t c) { return T[a] > T[c]; });
      if (!b->d_order) b->d_order = dalloc<int32_t>(b->capacity);
      b->tuned_count = -1;   // the tuned order is overwritten
      HPO_HIP(hipMemcpy(b->d_order, idx.data(), count * sizeof(int32_t), hipMemcpyHostToDevice));
      order = b->d_order;
    }
    hpo_timing local{};
    launch_integrate(m, count, b->d_y, b->d_y, b->ld, b->d_P, b->d_dt, b->d_status, b->d_stats, order, nullptr,
                     opt, st, &local);
    if (tm) *tm = local;
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

int hpo_batch_get(const hpo_batch* b, int64_t first, int64_t count, double* y, int64_t ldy, int32_t* status,
                  int32_t* stats) {
  if (!b) return fail(HPO_E_INVALID, "null batch");
  if (first < 0 || count < 0 || first + count > b->capacity) return fail(HPO_E_INVALID, "range exceeds capacity");
  if (count == 0) return HPO_OK;
  try {
    HPO_HIP(hipSetDevice(b->device));
    if (y) {
      if (ldy < b->ld) return fail(HPO_E_INVALID, "ldy smaller than the state length");
      HPO_HIP(hipMemcpy2D(y, ldy * sizeof(double), b->d_y + first * b->ld, b->ld * sizeof(double),
                          b->ld * sizeof(double), count, hipMemcpyDeviceToHost));
    }
    if (status)
      HPO_HIP(hipMemcpy(status, b->d_status + first, count * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (stats)
      HPO_HIP(hipMemcpy(stats, b->d_stats + first * HPO_NSTATS, count * HPO_NSTATS * sizeof(int32_t),
                        hipMemcpyDeviceToHost));
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

int hpo_integrate_device(const hpo_mech* mc, int64_t ncells, const double* d_y_in, double* d_y_out, int64_t ld,
                         const double* d_P, const double* d_dt, int32_t* d_status, int32_t* d_stats,
                         const int32_t* d_order, int64_t* d_prof, const hpo_opts* o, void* stream,
                         hpo_timing* tm) {
  if (!mc || !d_y_in || !d_y_out || !d_P || !d_dt || !d_status) return fail(HPO_E_INVALID, "null argument");
  hpo_opts def;
  hpo_opts_default(&def);
  try {
    launch_integrate(const_cast<hpo_mech*>(mc), ncells, d_y_in, d_y_out, ld, d_P, d_dt, d_status, d_stats, d_order,
                     d_prof, o ? *o : def, static_cast<hipStream_t>(stream), tm);
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

int hpo_integrate_device_noalloc(const hpo_mech* mc, int64_t ncells, const double* d_y_in, double* d_y_out,
                                 int64_t ld, const double* d_P, const double* d_dt, int32_t* d_status,
                                 int32_t* d_stats, const int32_t* d_order, const hpo_opts* o, void* stream) {
  if (!mc || !d_y_in || !d_y_out || !d_P || !d_dt || !d_status) return fail(HPO_E_INVALID, "null argument");
  hpo_opts def;
  hpo_opts_default(&def);
  const hpo_opts& opt = o ? *o : def;
  try {
    check_opts(opt);
    const hpo_mech* m = mc;
    const int ns = m->host.K();
    if (use_block_layout(m, opt))
      throw std::invalid_argument("the no-allocation entry point runs the wave-per-cell kernel only");
    if (ncells < 0 || ncells > 0x7fffffffLL) throw std::invalid_argument("ncells out of range");
    if (ld < ns) throw std::invalid_argument("leading dimension smaller than the state");
    int device = -1;
    HPO_HIP(hipGetDevice(&device));
    if (!m->d_blob || m->device != device)
      throw std::invalid_argument(
          "the mechanism's tables are not on the current device: run one allocating call first "
          "(hpo_integrate_device with ncells = 0 uploads them)");
    void* base = nullptr;
    HPO_HIP(hpo::noalloc_scratch(device, &base));
    const int wpb = opt.waves_per_block > 0 ? opt.waves_per_block : 4;
    const int block = 64 * wpb;
    const size_t lds = static_cast<size_t>(m->blob.lay.lds_bytes) +
                       static_cast<size_t>(wpb) * hpo::bdf_wave_lds_doubles(m->blob.lay) * 8;
    const int variant = hpo::bdf_variant(opt.lu_precision, opt.lu_layout);
    int occ = 0;
    HPO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, hpo::bdf_kernel_fn(ns, variant, 0, opt.reactor), block, lds));
    if (occ < 1) throw std::invalid_argument("kernel does not fit on a CU with this configuration");
    const int bpc = opt.blocks_per_cu > 0 ? std::min(opt.blocks_per_cu, occ) : occ;
    int64_t grid = static_cast<int64_t>(m->n_cu) * bpc;
    grid = std::max<int64_t>(std::min<int64_t>(grid, (ncells + wpb - 1) / wpb), 1);
    const int64_t stride = m->scratch_stride;
    const size_t need = 256 + static_cast<size_t>(grid * wpb * stride) * sizeof(double);
    if (need > hpo::kNoallocScratchBytes)
      throw std::invalid_argument("launch needs " + std::to_string(need) + " bytes of scratch, more than the " +
                                  std::to_string(hpo::kNoallocScratchBytes) +
                                  " bytes of the no-allocation entry point (fewer blocks_per_cu fit)");
    hpo::BdfArgs A{};
    A.mech = m->d_blob;
    A.lay = m->blob.lay;
    A.y_in = d_y_in;
    A.y_out = d_y_out;
    A.P = d_P;
    A.dt = d_dt;
    A.status = d_status;
    A.stats = d_stats;
    A.order = d_order;
    A.queue = static_cast<uint32_t*>(base);
    A.scratch = reinterpret_cast<double*>(static_cast<uint8_t*>(base) + 256);
    A.scratch_stride = stride;
    A.ncells = ncells;
    A.ld = ld;
    A.rtol = opt.rtol;
    A.atol = opt.atol;
    A.maxsteps = opt.maxsteps;
    A.refine_sweeps = opt.refine_sweeps;
    A.prof = nullptr;
    A.reactor = opt.reactor;
    // the FP64 fallback needs the input rows after the FP32 launch: not in place here, where
    // there is no memory to copy them to
    const int64_t rows = ncells * ld;
    const bool in_place = rows_overlap(d_y_in, d_y_out, rows);
    const bool fallback = opt.lu_precision == 1 && opt.fp64_fallback && ncells > 0 && !in_place;
    hipStream_t st = static_cast<hipStream_t>(stream);
    HPO_HIP(hpo::launch_reset_counter(A.queue, fallback, st));
    if (ncells > 0) {
      HPO_HIP(hpo::launch_bdf(ns, A, variant, static_cast<int>(grid), block, lds, st));
      if (fallback)
        launch_fp64_fallback(m, false, A, d_y_in, wpb, block, lds, 100,
                             static_cast<int64_t>((hpo::kNoallocScratchBytes - 256) / sizeof(double)), st);
    }
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

// cell_layout: 0 = the wave tables when the mechanism has them, else the block tables;
// 1 = wave (one wave per cell); 2 = block (one 256-thread workgroup per cell)
static int eval_common(const hpo_mech* mc, int32_t cell_layout, int32_t reactor, int64_t n, const double* y,
                       const double* P, double* out, int mode) {
  if (!mc || !y || !P || !out || n < 0) return fail(HPO_E_INVALID, "bad argument");
  if (cell_layout < 0 || cell_layout > 2) return fail(HPO_E_INVALID, "cell_layout must be 0, 1 or 2");
  if (reactor < 0 || reactor > 1) return fail(HPO_E_INVALID, "reactor must be 0 (constant pressure) or 1 (constant volume)");
  hpo_mech* m = const_cast<hpo_mech*>(mc);
  const int ns = m->host.K();
  const bool wave_ok = m->has_wave && hpo::kernel_supports(ns);
  const bool block_ok = m->has_block && hpo::block_kernel_supports(ns);
  const bool block = cell_layout == 2 || (cell_layout == 0 && !wave_ok);
  if (block ? !block_ok : !wave_ok) return fail(HPO_E_UNSUPPORTED, "no kernel of this layout for this state size");
  if (reactor == 1 && m->has_block && m->bblob.lay.n_plog > 0)
    return fail(HPO_E_INVALID, "PLOG rates are supported at constant pressure only");
  if (n == 0) return HPO_OK;
  double *dy = nullptr, *dP = nullptr, *dout = nullptr;
  int rc = HPO_OK;
  try {
    m->ensure_device();
    // the handle's scratch: after any integration still running on it
    std::lock_guard<std::mutex> g(m->launch_mu);
    if (m->launched) HPO_HIP(hipEventSynchronize(m->ev_last));
    const int wpb = block ? 1 : 4;
    const int64_t grid = std::min<int64_t>((n + wpb - 1) / wpb, block ? 1024 : 4096);
    const int64_t stride = block ? hpo::blk_scratch_doubles_host(m->bblob.lay) : m->scratch_stride;
    m->ensure_scratch_doubles(grid * wpb * stride);
    size_t outn = mode == 0 ? static_cast<size_t>(n * ns) : static_cast<size_t>(n * ns * ns);
    dy = dalloc<double>(static_cast<size_t>(n * ns));
    dP = dalloc<double>(n);
    dout = dalloc<double>(outn);
    HPO_HIP(hipMemcpy(dy, y, n * ns * sizeof(double), hipMemcpyHostToDevice));
    HPO_HIP(hipMemcpy(dP, P, n * sizeof(double), hipMemcpyHostToDevice));
    hpo::EvalArgs A{};
    A.mech = block ? m->d_bblob : m->d_blob;
    A.lay = m->blob.lay;
    A.blay = m->bblob.lay;
    A.y = dy;
    A.P = dP;
    A.out = dout;
    A.scratch = m->d_scratch;
    A.scratch_stride = stride;
    A.ncells = n;
    A.ld = ns;
    A.mode = mode;
    A.reactor = reactor;
    if (block) {
      const size_t lds = static_cast<size_t>(m->bblob.lay.lds_bytes) + hpo::blk_work_lds_bytes(m->bblob.lay);
      HPO_HIP(hpo::launch_eval_block(ns, A, static_cast<int>(grid), lds, nullptr));
    } else {
      const size_t lds = static_cast<size_t>(m->blob.lay.lds_bytes) + wpb * hpo::eval_wave_lds_doubles(m->blob.lay) * 8;
      HPO_HIP(hpo::launch_eval(ns, A, static_cast<int>(grid), 64 * wpb, lds, nullptr));
    }
    HPO_HIP(hipDeviceSynchronize());
    HPO_HIP(hipMemcpy(out, dout, outn * sizeof(double), hipMemcpyDeviceToHost));
  } catch (...) {
    rc = map_exceptions();
  }
  if (dy) (void)hipFree(dy);
  if (dP) (void)hipFree(dP);
  if (dout) (void)hipFree(dout);
  return rc;
}

int hpo_eval_rhs(const hpo_mech* m, int64_t n, const double* y, const double* P, double* f) {
  return eval_common(m, 0, 0, n, y, P, f, 0);
}

int hpo_eval_jac(const hpo_mech* m, int64_t n, const double* y, const double* P, double* J) {
  return eval_common(m, 0, 0, n, y, P, J, 1);
}

int hpo_eval_layout(const hpo_mech* m, int32_t cell_layout, int32_t reactor, int32_t what, int64_t n,
                    const double* y, const double* P, double* out) {
  if (what != 0 && what != 1) return fail(HPO_E_INVALID, "what must be 0 (RHS) or 1 (Jacobian)");
  return eval_common(m, cell_layout, reactor, n, y, P, out, what);
}

int hpo_lu_solve(int32_t ns, int32_t fp32, int64_t n, const double* M, const double* b, double* x, int32_t* ok) {
  if (!M || !b || !x || !ok || n < 0) return fail(HPO_E_INVALID, "bad argument");
  if (!hpo::kernel_supports(ns)) return fail(HPO_E_UNSUPPORTED, "no kernel for this size");
  if (n == 0) return HPO_OK;
  double *dM = nullptr, *db = nullptr, *dx = nullptr;
  int32_t* dok = nullptr;
  int rc = HPO_OK;
  try {
    dM = dalloc<double>(static_cast<size_t>(n * ns * ns));
    db = dalloc<double>(static_cast<size_t>(n * ns));
    dx = dalloc<double>(static_cast<size_t>(n * ns));
    dok = dalloc<int32_t>(n);
    HPO_HIP(hipMemcpy(dM, M, n * ns * ns * sizeof(double), hipMemcpyHostToDevice));
    HPO_HIP(hipMemcpy(db, b, n * ns * sizeof(double), hipMemcpyHostToDevice));
    hpo::LuArgs A{dM, db, dx, dok};
    HPO_HIP(hpo::launch_lu(ns, fp32, A, n, nullptr));
    HPO_HIP(hipDeviceSynchronize());
    HPO_HIP(hipMemcpy(x, dx, n * ns * sizeof(double), hipMemcpyDeviceToHost));
    HPO_HIP(hipMemcpy(ok, dok, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  } catch (...) {
    rc = map_exceptions();
  }
  for (void* p : {static_cast<void*>(dM), static_cast<void*>(db), static_cast<void*>(dx), static_cast<void*>(dok)})
    if (p) (void)hipFree(p);
  return rc;
}

int hpo_lu_solve_layout(int32_t ns, int32_t fp32, int32_t layout, int64_t n, const double* M, const double* b,
                        double* x, int32_t* ok, float* kernel_ms) {
  if (!M || !b || !x || !ok || n < 0) return fail(HPO_E_INVALID, "bad argument");
  if (layout < 0 || layout > 2)
    return fail(HPO_E_INVALID, "layout must be 0 (wave per matrix), 1 (workgroup LU) or 2 (workgroup Gauss-Jordan)");
  if (layout == 1 && fp32) return fail(HPO_E_UNSUPPORTED, "the workgroup LU layout is FP64 only");
  if (layout == 2 ? !hpo::block_kernel_supports(ns) : !hpo::kernel_supports(ns))
    return fail(HPO_E_UNSUPPORTED, "no kernel for this size");
  if (n > 0x7fffffffLL) return fail(HPO_E_INVALID, "too many matrices for one launch");
  if (kernel_ms) *kernel_ms = 0.f;
  if (n == 0) return HPO_OK;
  double *dM = nullptr, *db = nullptr, *dx = nullptr;
  int32_t* dok = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = HPO_OK;
  try {
    dM = dalloc<double>(static_cast<size_t>(n * ns * ns));
    db = dalloc<double>(static_cast<size_t>(n * ns));
    dx = dalloc<double>(static_cast<size_t>(n * ns));
    dok = dalloc<int32_t>(n);
    HPO_HIP(hipMemcpy(dM, M, n * ns * ns * sizeof(double), hipMemcpyHostToDevice));
    HPO_HIP(hipMemcpy(db, b, n * ns * sizeof(double), hipMemcpyHostToDevice));
    HPO_HIP(hipEventCreate(&e0));
    HPO_HIP(hipEventCreate(&e1));
    hpo::LuArgs A{dM, db, dx, dok};
    HPO_HIP(hipEventRecord(e0, nullptr));
    if (layout == 0)
      HPO_HIP(hpo::launch_lu(ns, fp32, A, n, nullptr));
    else if (layout == 1)
      HPO_HIP(hpo::launch_lu_block(ns, A, n, nullptr));
    else
      HPO_HIP(hpo::launch_gj_block(ns, fp32, A, n, nullptr));
    HPO_HIP(hipEventRecord(e1, nullptr));
    HPO_HIP(hipEventSynchronize(e1));
    if (kernel_ms) HPO_HIP(hipEventElapsedTime(kernel_ms, e0, e1));
    HPO_HIP(hipMemcpy(x, dx, n * ns * sizeof(double), hipMemcpyDeviceToHost));
    HPO_HIP(hipMemcpy(ok, dok, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  } catch (...) {
    rc = map_exceptions();
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* p : {static_cast<void*>(dM), static_cast<void*>(db), static_cast<void*>(dx), static_cast<void*>(dok)})
    if (p) (void)hipFree(p);
  return rc;
}

namespace {

// Mean step count per (T0, log10 dt) bin of the calibration sample; bins the sample left
// empty take their temperature row's mean, then the overall mean.
struct CostModel {
  static constexpr int NT = 24, ND = 12;
  double t0 = 0, t1 = 1, d0 = 0, d1 = 1;
  double mean[NT][ND];
  int bin_t(double T) const { return std::min(NT - 1, std::max(0, static_cast<int>((T - t0) / (t1 - t0) * NT))); }
  int bin_d(double dt) const {
    double l = std::log10(std::max(dt, 1e-300));
    return std::min(ND - 1, std::max(0, static_cast<int>((l - d0) / (d1 - d0) * ND)));
  }
  void fit(const std::vector<double>& T, const std::vector<double>& dt, const std::vector<double>& cost) {
    t0 = *std::min_element(T.begin(), T.end());
    t1 = *std::max_element(T.begin(), T.end()) + 1e-9;
    d0 = std::log10(std::max(*std::min_element(dt.begin(), dt.end()), 1e-300));
    d1 = std::log10(std::max(*std::max_element(dt.begin(), dt.end()), 1e-300)) + 1e-9;
    double sum[NT][ND] = {}, cnt[NT][ND] = {};
    double all = 0;
    for (size_t i = 0; i < T.size(); ++i) {
      sum[bin_t(T[i])][bin_d(dt[i])] += cost[i];
      cnt[bin_t(T[i])][bin_d(dt[i])] += 1;
      all += cost[i];
    }
    all /= std::max<size_t>(T.size(), 1);
    for (int a = 0; a < NT; ++a) {
      double rs = 0, rc = 0;
      for (int b = 0; b < ND; ++b) rs += sum[a][b], rc += cnt[a][b];
      for (int b = 0; b < ND; ++b)
        mean[a][b] = cnt[a][b] > 0 ? sum[a][b] / cnt[a][b] : (rc > 0 ? rs / rc : all);
    }
  }
  double operator()(double T, double dt) const { return mean[bin_t(T)][bin_d(dt)]; }
};

// Order of cells by a key, descending, ties by index.
std::vector<int32_t> order_desc(const std::vector<double>& key) {
  std::vector<int32_t> idx(key.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return key[a] > key[b]; });
  return idx;
}

}  // namespace

int hpo_autotune(const hpo_mech* mc, int64_t ncells, const double* d_y, int64_t ld, const double* d_P,
                 const double* d_dt, int64_t sample_cells, hpo_opts* o, int32_t* d_order_out,
                 hpo_tune_report* rep, void* stream) {
  if (!mc || !d_y || !d_P || !d_dt || !o || !d_order_out || ncells <= 0) return fail(HPO_E_INVALID, "bad argument");
  hpo_mech* m = const_cast<hpo_mech*>(mc);
  const auto wall0 = std::chrono::steady_clock::now();
  hipStream_t st = static_cast<hipStream_t>(stream);
  double *sy = nullptr, *sP = nullptr, *sdt = nullptr, *syo = nullptr, *dT0 = nullptr;
  int32_t *sst = nullptr, *sstats = nullptr, *sorder = nullptr;
  int rc = HPO_OK;
  try {
    check_opts(*o);
    m->ensure_device();
    const int64_t ns = m->host.K();
    if (ld < ns) throw std::invalid_argument("leading dimension smaller than the state");
    // evenly strided sample
    const int64_t ns_req = sample_cells > 0 ? sample_cells : 32768;
    const int64_t stride = std::max<int64_t>(1, ncells / std::max<int64_t>(1, ns_req));
    const int64_t S = std::min<int64_t>(ns_req, (ncells + stride - 1) / stride);
    sy = dalloc<double>(static_cast<size_t>(S * ns));
    syo = dalloc<double>(static_cast<size_t>(S * ns));
    sP = dalloc<double>(S);
    sdt = dalloc<double>(S);
    sst = dalloc<int32_t>(S);
    sstats = dalloc<int32_t>(static_cast<size_t>(S * HPO_NSTATS));
    sorder = dalloc<int32_t>(S);
    HPO_HIP(hipMemcpy2DAsync(sy, ns * sizeof(double), d_y, stride * ld * sizeof(double), ns * sizeof(double), S,
                             hipMemcpyDeviceToDevice, st));
    HPO_HIP(hipMemcpy2DAsync(sP, sizeof(double), d_P, stride * sizeof(double), sizeof(double), S,
                             hipMemcpyDeviceToDevice, st));
    HPO_HIP(hipMemcpy2DAsync(sdt, sizeof(double), d_dt, stride * sizeof(double), sizeof(double), S,
                             hipMemcpyDeviceToDevice, st));
    // T0 and dt of every cell (host), for the orders of the full workload
    // (the strided temperature column is packed on the device first: a strided copy straight
    // to the host moves one 8-byte row per transfer)
    std::vector<double> T(ncells), dt(ncells);
    dT0 = dalloc<double>(ncells);
    HPO_HIP(hipMemcpy2DAsync(dT0, sizeof(double), d_y, ld * sizeof(double), sizeof(double), ncells,
                             hipMemcpyDeviceToDevice, st));
    HPO_HIP(hipMemcpyAsync(T.data(), dT0, ncells * sizeof(double), hipMemcpyDeviceToHost, st));
    HPO_HIP(hipMemcpyAsync(dt.data(), d_dt, ncells * sizeof(double), hipMemcpyDeviceToHost, st));
    HPO_HIP(hipStreamSynchronize(st));
    std::vector<double> sT(S), sdtv(S);
    for (int64_t i = 0; i < S; ++i) sT[i] = T[i * stride], sdtv[i] = dt[i * stride];

    // calibration run: step counts of the sample cells
    hpo_opts base = *o;
    base.waves_per_block = 0;
    base.blocks_per_cu = 0;
    base.sort_cells = 0;
    hpo_timing tm{};
    launch_integrate(m, S, sy, syo, ns, sP, sdt, sst, sstats, nullptr, nullptr, base, st, &tm);
    std::vector<int32_t> stats(static_cast<size_t>(S * HPO_NSTATS));
    HPO_HIP(hipMemcpy(stats.data(), sstats, stats.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<double> cost(S);
    for (int64_t i = 0; i < S; ++i) cost[i] = stats[i * HPO_NSTATS + 1];   // nfe: RHS evaluations
    CostModel cm;
    cm.fit(sT, sdtv, cost);

    // sample orders for the three modes
    std::vector<std::vector<int32_t>> sample_orders(3);
    sample_orders[0].resize(S);
    std::iota(sample_orders[0].begin(), sample_orders[0].end(), 0);
    sample_orders[1] = order_desc(sT);
    std::vector<double> pred(S);
    for (int64_t i = 0; i < S; ++i) pred[i] = cm(sT[i], sdtv[i]);
    sample_orders[2] = order_desc(pred);

    hpo_tune_report R{};
    R.sample_cells = S;
    const bool block_only = use_block_layout(m, base);
    const bool block_too = !block_only && o->cell_layout == 0 && m->has_block && hpo::block_kernel_supports(ns);
    const int wpbs[3] = {4, 2, 1};
    double best_ms = 1e300;
    int best = -1;
    auto record = [&](const hpo_opts& c, int om, const hpo_timing& t) {
      const int k = R.n_candidates++;
      const bool blk = use_block_layout(m, c);
      R.waves_per_block[k] = blk ? 4 : c.waves_per_block;
      R.blocks_per_cu[k] = (t.grid > 0 && m->n_cu > 0) ? std::max(1, t.grid / m->n_cu) : 0;
      R.order_mode[k] = om;
      R.lu_layout[k] = c.lu_layout;
      R.cell_layout[k] = blk ? 2 : 1;
      R.sample_ms[k] = t.kernel_ms;
      if (t.kernel_ms < best_ms) best_ms = t.kernel_ms, best = k;
    };
    for (int wi = 0; wi < (block_only ? 1 : 3); ++wi)
      for (int om = 0; om < 3 && R.n_candidates < HPO_TUNE_MAX; ++om) {
        HPO_HIP(hipMemcpyAsync(sorder, sample_orders[om].data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
        hpo_opts c = base;
        c.waves_per_block = wpbs[wi];
        hpo_timing t{};
        launch_integrate(m, S, sy, syo, ns, sP, sdt, sst, nullptr, sorder, nullptr, c, st, &t);
        record(c, om, t);
      }
    // one workgroup per cell (the block layout) at the best order, where it exists
    if (block_too && best >= 0 && R.n_candidates < HPO_TUNE_MAX) {
      const int om = R.order_mode[best];
      HPO_HIP(hipMemcpyAsync(sorder, sample_orders[om].data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
      hpo_opts c = base;
      c.cell_layout = 2;
      hpo_timing t{};
      launch_integrate(m, S, sy, syo, ns, sP, sdt, sst, nullptr, sorder, nullptr, c, st, &t);
      record(c, om, t);
    }
    // FP32 path: the other Newton-matrix layout (explicit inverse / in-place LU) at the best
    // launch shape and order
    if (o->lu_precision == 1 && best >= 0 && R.cell_layout[best] == 1 && R.n_candidates < HPO_TUNE_MAX) {
      const int om = R.order_mode[best];
      HPO_HIP(hipMemcpyAsync(sorder, sample_orders[om].data(), S * sizeof(int32_t), hipMemcpyHostToDevice, st));
      hpo_opts c = base;
      c.waves_per_block = R.waves_per_block[best];
      c.lu_layout = 1 - base.lu_layout;
      c.cell_layout = 1;
      hpo_timing t{};
      launch_integrate(m, S, sy, syo, ns, sP, sdt, sst, nullptr, sorder, nullptr, c, st, &t);
      record(c, om, t);
    }
    R.best = best;
    o->waves_per_block = R.waves_per_block[best];
    o->lu_layout = R.lu_layout[best];
    o->cell_layout = R.cell_layout[best];
    o->blocks_per_cu = 0;   // as many workgroups per CU as fit (what every candidate ran with)
    o->sort_cells = 0;

    // the chosen order over the whole workload
    std::vector<int32_t> full;
    if (R.order_mode[best] == 0) {
      full.resize(ncells);
      std::iota(full.begin(), full.end(), 0);
    } else if (R.order_mode[best] == 1) {
      full = order_desc(T);
    } else {
      std::vector<double> p(ncells);
      for (int64_t i = 0; i < ncells; ++i) p[i] = cm(T[i], dt[i]);
      full = order_desc(p);
    }
    HPO_HIP(hipMemcpyAsync(d_order_out, full.data(), ncells * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HPO_HIP(hipStreamSynchronize(st));
    R.tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    if (rep) *rep = R;
  } catch (...) {
    rc = map_exceptions();
  }
  (void)hipStreamSynchronize(st);
  for (void* p : {static_cast<void*>(sy), static_cast<void*>(syo), static_cast<void*>(sP), static_cast<void*>(sdt),
                  static_cast<void*>(sst), static_cast<void*>(sstats), static_cast<void*>(sorder),
                  static_cast<void*>(dT0)})
    if (p) (void)hipFree(p);
  return rc;
}

int hpo_batch_autotune(hpo_batch* b, int64_t count, int64_t sample_cells, hpo_opts* o, hpo_tune_report* rep,
                       void* stream) {
  if (!b || !o) return fail(HPO_E_INVALID, "null argument");
  if (count <= 0 || count > b->capacity) return fail(HPO_E_INVALID, "count out of range");
  try {
    HPO_HIP(hipSetDevice(b->device));
    if (!b->d_order) b->d_order = dalloc<int32_t>(b->capacity);
  } catch (...) {
    return map_exceptions();
  }
  b->tuned_count = -1;
  int rc = hpo_autotune(b->mech, count, b->d_y, b->ld, b->d_P, b->d_dt, sample_cells, o, b->d_order, rep, stream);
  if (rc == HPO_OK) {
    b->tuned_count = count;
    o->sort_cells = 2;
  }
  return rc;
}

int hpo_device_info(int32_t* n_cu, int64_t* free_bytes, int64_t* total_bytes) {
  try {
    int dev = 0;
    HPO_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HPO_HIP(hipGetDeviceProperties(&prop, dev));
    if (n_cu) *n_cu = prop.multiProcessorCount;
    size_t fr = 0, to = 0;
    HPO_HIP(hipMemGetInfo(&fr, &to));
    if (free_bytes) *free_bytes = static_cast<int64_t>(fr);
    if (total_bytes) *total_bytes = static_cast<int64_t>(to);
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

}  // extern "C"
