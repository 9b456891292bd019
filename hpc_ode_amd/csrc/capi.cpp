// C ABI implementation (include/hpc_ode.h): handles, device memory, launches.
#include "../../include/hpc_ode.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "mechanism.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct HipError : std::runtime_error {
  hipError_t e;
  HipError(hipError_t err, const std::string& what) : std::runtime_error(what), e(err) {}
};

#define HPO_HIP(call)                                                                               \
  do {                                                                                              \
    hipError_t _e = (call);                                                                         \
    if (_e != hipSuccess)                                                                           \
      throw HipError(_e, std::string(#call) + ": " + hipGetErrorString(_e));                        \
  } while (0)

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

template <typename T>
T* dalloc(size_t n) {
  void* p = nullptr;
  HPO_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

}  // namespace

struct hpo_mech {
  hpo::Mechanism host;
  hpo::DevMechBlob blob;
  int device = -1;
  int n_cu = 0;
  uint8_t* d_blob = nullptr;
  double* d_scratch = nullptr;
  int64_t scratch_slots = 0;
  int64_t scratch_stride = 0;
  uint32_t* d_queue = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  ~hpo_mech() {
    if (device >= 0) {
      (void)hipSetDevice(device);
      if (d_blob) (void)hipFree(d_blob);
      if (d_scratch) (void)hipFree(d_scratch);
      if (d_queue) (void)hipFree(d_queue);
      if (ev0) (void)hipEventDestroy(ev0);
      if (ev1) (void)hipEventDestroy(ev1);
    }
  }

  // Uploads the tables on first use (on the current device).
  void ensure_device() {
    if (d_blob) return;
    HPO_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    HPO_HIP(hipGetDeviceProperties(&prop, device));
    n_cu = prop.multiProcessorCount;
    d_blob = dalloc<uint8_t>(blob.bytes.size());
    HPO_HIP(hipMemcpy(d_blob, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice));
    d_queue = dalloc<uint32_t>(4);
    HPO_HIP(hipEventCreate(&ev0));
    HPO_HIP(hipEventCreate(&ev1));
    scratch_stride = hpo::wave_scratch_doubles_host(blob.lay);
  }

  void ensure_scratch(int64_t slots) {
    if (slots <= scratch_slots) return;
    if (d_scratch) HPO_HIP(hipFree(d_scratch));
    d_scratch = nullptr;
    d_scratch = dalloc<double>(static_cast<size_t>(slots * scratch_stride));
    scratch_slots = slots;
  }
};

struct hpo_batch {
  const hpo_mech* mech = nullptr;
  int64_t capacity = 0;
  int64_t ld = 0;
  int device = -1;
  double *d_y = nullptr, *d_P = nullptr, *d_dt = nullptr;
  int32_t *d_status = nullptr, *d_stats = nullptr, *d_order = nullptr;
  ~hpo_batch() {
    if (device >= 0) {
      (void)hipSetDevice(device);
      for (void* p : {static_cast<void*>(d_y), static_cast<void*>(d_P), static_cast<void*>(d_dt),
                      static_cast<void*>(d_status), static_cast<void*>(d_stats), static_cast<void*>(d_order)})
        if (p) (void)hipFree(p);
    }
  }
};

namespace {

void check_opts(const hpo_opts& o) {
  if (!(o.rtol >= 0.0) || !(o.atol >= 0.0)) throw std::invalid_argument("tolerances must be >= 0");
  if (o.lu_precision != 0 && o.lu_precision != 1) throw std::invalid_argument("lu_precision must be 0 or 1");
  if (o.refine_sweeps < 0 || o.refine_sweeps > 8) throw std::invalid_argument("refine_sweeps must be in 0..8");
  if (o.waves_per_block < 0 || o.waves_per_block > 4) throw std::invalid_argument("waves_per_block must be in 0..4");
  if (o.blocks_per_cu < 0) throw std::invalid_argument("blocks_per_cu must be >= 0");
}

// Launches the persistent BDF kernel on arrays already resident on the device.
void launch_integrate(hpo_mech* m, int64_t ncells, const double* d_yin, double* d_yout, int64_t ld,
                      const double* d_P, const double* d_dt, int32_t* d_status, int32_t* d_stats,
                      const int32_t* d_order, int64_t* d_prof, const hpo_opts& o, hipStream_t st,
                      hpo_timing* tm) {
  check_opts(o);
  const int ns = m->blob.lay.K;
  if (!hpo::kernel_supports(ns))
    throw std::invalid_argument("no kernel instantiated for a state of " + std::to_string(ns) + " variables");
  if (ncells < 0 || ncells > 0x7fffffffLL) throw std::invalid_argument("ncells out of range");
  if (ld < ns) throw std::invalid_argument("leading dimension smaller than the state");
  m->ensure_device();
  const int wpb = o.waves_per_block > 0 ? o.waves_per_block : 4;
  const int block = 64 * wpb;
  const size_t lds = static_cast<size_t>(m->blob.lay.lds_bytes) +
                     static_cast<size_t>(wpb) * hpo::bdf_wave_lds_doubles(m->blob.lay) * 8;
  const void* fn = hpo::bdf_kernel_fn(ns, o.lu_precision);
  int occ = 0;
  HPO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, block, lds));
  if (occ < 1) throw std::invalid_argument("kernel does not fit on a CU with this configuration");
  int bpc = o.blocks_per_cu > 0 ? std::min(o.blocks_per_cu, occ) : occ;
  int64_t grid = static_cast<int64_t>(m->n_cu) * bpc;
  grid = std::min<int64_t>(grid, (ncells + wpb - 1) / wpb);
  grid = std::max<int64_t>(grid, 1);
  m->ensure_scratch(grid * wpb);

  hpo::BdfArgs A{};
  A.mech = m->d_blob;
  A.lay = m->blob.lay;
  A.y_in = d_yin;
  A.y_out = d_yout;
  A.P = d_P;
  A.dt = d_dt;
  A.status = d_status;
  A.stats = d_stats;
  A.order = d_order;
  A.queue = m->d_queue;
  A.scratch = m->d_scratch;
  A.scratch_stride = m->scratch_stride;
  A.ncells = ncells;
  A.ld = ld;
  A.rtol = o.rtol;
  A.atol = o.atol;
  A.maxsteps = o.maxsteps;
  A.refine_sweeps = o.refine_sweeps;
  A.prof = reinterpret_cast<int64_t*>(d_prof);

  HPO_HIP(hipMemsetAsync(m->d_queue, 0, sizeof(uint32_t), st));
  HPO_HIP(hipEventRecord(m->ev0, st));
  if (ncells > 0) HPO_HIP(hpo::launch_bdf(ns, A, o.lu_precision, static_cast<int>(grid), block, lds, st));
  HPO_HIP(hipEventRecord(m->ev1, st));
  if (tm) {
    HPO_HIP(hipEventSynchronize(m->ev1));
    float ms = 0.f;
    HPO_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    tm->kernel_ms = ms;
    tm->ncells = ncells;
    tm->grid = static_cast<int32_t>(grid);
    tm->block = block;
    tm->lds_bytes = static_cast<int32_t>(lds);
  }
}

}  // namespace

extern "C" {

void hpo_opts_default(hpo_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->rtol = 1.0e-6;
  o->atol = 1.0e-10;
  o->maxsteps = 10000;
  o->lu_precision = 0;
  o->refine_sweeps = 1;
}

const char* hpo_last_error(void) { return g_err.c_str(); }

int hpo_mech_load_yaml(const char* path, hpo_mech** out) {
  if (!path || !out) return fail(HPO_E_INVALID, "null argument");
  *out = nullptr;
  try {
    auto m = std::make_unique<hpo_mech>();
    m->host = hpo::load_mechanism_yaml(path);
    m->blob = hpo::build_device_tables(m->host);
    *out = m.release();
    return HPO_OK;
  } catch (...) {
    return map_exceptions();
  }
}

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
    if (opt.sort_cells && count > 1) {
      // hottest cells first: they take the most steps, so the queue tail holds short cells
      std::vector<double> T(count);
      HPO_HIP(hipMemcpy2D(T.data(), sizeof(double), b->d_y, b->ld * sizeof(double), sizeof(double), count,
                          hipMemcpyDeviceToHost));
      std::vector<int32_t> idx(count);
      std::iota(idx.begin(), idx.end(), 0);
      std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t c) { return T[a] > T[c]; });
      if (!b->d_order) b->d_order = dalloc<int32_t>(b->capacity);
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

static int eval_common(const hpo_mech* mc, int64_t n, const double* y, const double* P, double* out, int mode) {
  if (!mc || !y || !P || !out || n < 0) return fail(HPO_E_INVALID, "bad argument");
  if (n == 0) return HPO_OK;
  hpo_mech* m = const_cast<hpo_mech*>(mc);
  const int ns = m->blob.lay.K;
  if (!hpo::kernel_supports(ns)) return fail(HPO_E_UNSUPPORTED, "no kernel for this state size");
  double *dy = nullptr, *dP = nullptr, *dout = nullptr;
  int rc = HPO_OK;
  try {
    m->ensure_device();
    const int wpb = 4, block = 256;
    int64_t grid = std::min<int64_t>((n + wpb - 1) / wpb, 4096);
    m->ensure_scratch(grid * wpb);
    size_t outn = mode == 0 ? static_cast<size_t>(n * ns) : static_cast<size_t>(n * ns * ns);
    dy = dalloc<double>(static_cast<size_t>(n * ns));
    dP = dalloc<double>(n);
    dout = dalloc<double>(outn);
    HPO_HIP(hipMemcpy(dy, y, n * ns * sizeof(double), hipMemcpyHostToDevice));
    HPO_HIP(hipMemcpy(dP, P, n * sizeof(double), hipMemcpyHostToDevice));
    hpo::EvalArgs A{};
    A.mech = m->d_blob;
    A.lay = m->blob.lay;
    A.y = dy;
    A.P = dP;
    A.out = dout;
    A.scratch = m->d_scratch;
    A.scratch_stride = m->scratch_stride;
    A.ncells = n;
    A.ld = ns;
    A.mode = mode;
    size_t lds = static_cast<size_t>(m->blob.lay.lds_bytes) + wpb * hpo::eval_wave_lds_doubles(m->blob.lay) * 8;
    HPO_HIP(hpo::launch_eval(ns, A, static_cast<int>(grid), block, lds, nullptr));
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
  return eval_common(m, n, y, P, f, 0);
}

int hpo_eval_jac(const hpo_mech* m, int64_t n, const double* y, const double* P, double* J) {
  return eval_common(m, n, y, P, J, 1);
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
