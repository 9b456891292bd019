"""hpc_ode_amd: MI355X-native batched stiff-chemistry BDF integration.

The product is the C ABI in include/hpc_ode.h, implemented by libhpc_ode.so
(hand-written HIP kernels for gfx950 plus the C++ host code). This module is a
thin ctypes binding of that ABI for Python callers (tests, benchmarks, the
serving loop). It never falls back to a CPU path: if the library is missing it
raises, and every call goes to the GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HPO_LIB") or os.path.join(HERE, "libhpc_ode.so")
NSTATS = 8
NPROF = 5
PROF_FIELDS = ("rhs", "jac", "lu", "solve", "cell")
STATS_FIELDS = ("nst", "nfe", "nje", "nsetups", "nni", "ncfn", "netf", "qlast")

CELL_STATUS = {0: "SUCCESS", -1: "TOO_MUCH_WORK", -2: "TOO_MUCH_ACC", -3: "ERR_FAILURE",
               -4: "CONV_FAILURE", -22: "ILL_INPUT", -27: "TOO_CLOSE"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class HpoError(RuntimeError):
    pass


class Opts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("maxsteps", C.c_int32),
                ("lu_precision", C.c_int32), ("refine_sweeps", C.c_int32),
                ("waves_per_block", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("sort_cells", C.c_int32), ("lu_layout", C.c_int32), ("cell_layout", C.c_int32),
                ("reactor", C.c_int32), ("fp64_fallback", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("ncells", C.c_int64),
                ("grid", C.c_int32), ("block", C.c_int32), ("lds_bytes", C.c_int32),
                ("fallback_cells", C.c_int32)]


TUNE_MAX = 16
ORDER_MODES = {0: "cell index", 1: "T0 descending", 2: "predicted cost descending"}
LU_LAYOUTS = {0: "inverse", 1: "lu"}


class TuneReport(C.Structure):
    _fields_ = [("n_candidates", C.c_int32), ("best", C.c_int32),
                ("waves_per_block", C.c_int32 * TUNE_MAX), ("blocks_per_cu", C.c_int32 * TUNE_MAX),
                ("order_mode", C.c_int32 * TUNE_MAX), ("lu_layout", C.c_int32 * TUNE_MAX),
                ("cell_layout", C.c_int32 * TUNE_MAX), ("sample_ms", C.c_double * TUNE_MAX),
                ("sample_cells", C.c_int64), ("tune_ms", C.c_double)]

    def candidates(self):
        return [dict(waves_per_block=self.waves_per_block[i], blocks_per_cu=self.blocks_per_cu[i],
                     order=ORDER_MODES[self.order_mode[i]], lu_layout=LU_LAYOUTS[self.lu_layout[i]],
                     cell_layout=CELL_LAYOUTS[self.cell_layout[i]], sample_ms=self.sample_ms[i])
                for i in range(self.n_candidates)]


class ServerOpts(C.Structure):
    _fields_ = [("n_slots", C.c_int32), ("cu_share_pct", C.c_int32), ("max_batch_cells", C.c_int64),
                ("hbm_budget_bytes", C.c_int64), ("opts", Opts), ("launch_cells", C.c_int64)]


class RequestInfo(C.Structure):
    _fields_ = [("latency_ms", C.c_double), ("queue_ms", C.c_double), ("run_ms", C.c_double),
                ("ncells", C.c_int64), ("batch_cells", C.c_int64), ("batch_requests", C.c_int32),
                ("failed_cells", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


E_INVALID = -1
E_BUSY = -6
E_TIMEOUT = -7

_lib = None


def lib():
    """Loads libhpc_ode.so (building it first if it does not exist)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.hpo_last_error.restype = C.c_char_p
        L.hpo_opts_default.argtypes = [C.POINTER(Opts)]
        L.hpo_mech_load_yaml.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.hpo_mech_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.hpo_mech_destroy.argtypes = [vp]
        L.hpo_mech_info.argtypes = [vp, _ip, _ip, _ip]
        L.hpo_mech_species_name.argtypes = [vp, C.c_int32, C.c_char_p, C.c_int32]
        L.hpo_mech_molecular_weights.argtypes = [vp, _dp]
        L.hpo_mech_tables.argtypes = [vp, _dp, _dp, _dp, _ip, _ip]
        L.hpo_batch_create.argtypes = [vp, C.c_int64, C.POINTER(vp)]
        L.hpo_batch_destroy.argtypes = [vp]
        L.hpo_batch_put.argtypes = [vp, C.c_int64, C.c_int64, _dp, C.c_int64, _dp, C.c_double, _dp,
                                    C.c_double]
        L.hpo_batch_integrate.argtypes = [vp, C.c_int64, C.POINTER(Opts), vp, C.POINTER(Timing)]
        L.hpo_batch_get.argtypes = [vp, C.c_int64, C.c_int64, _dp, C.c_int64, _ip, _ip]
        L.hpo_integrate_device.argtypes = [vp, C.c_int64, vp, vp, C.c_int64, vp, vp, vp, vp, vp, vp,
                                           C.POINTER(Opts), vp, C.POINTER(Timing)]
        L.hpo_integrate_device_noalloc.argtypes = [vp, C.c_int64, vp, vp, C.c_int64, vp, vp, vp, vp, vp,
                                                   C.POINTER(Opts), vp]
        L.hpo_mech_enthalpy_mass.argtypes = [vp, C.c_int64, _dp, _dp, C.c_int64, _dp]
        L.hpo_eval_rhs.argtypes = [vp, C.c_int64, _dp, _dp, _dp]
        L.hpo_eval_jac.argtypes = [vp, C.c_int64, _dp, _dp, _dp]
        L.hpo_eval_layout.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int64, _dp, _dp, _dp]
        L.hpo_lu_solve.argtypes = [C.c_int32, C.c_int32, C.c_int64, _dp, _dp, _dp, _ip]
        L.hpo_lu_solve_layout.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int64, _dp, _dp, _dp, _ip,
                                          C.POINTER(C.c_float)]
        L.hpo_device_info.argtypes = [_ip, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.hpo_autotune.argtypes = [vp, C.c_int64, vp, C.c_int64, vp, vp, C.c_int64, C.POINTER(Opts), vp,
                                   C.POINTER(TuneReport), vp]
        L.hpo_batch_autotune.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(Opts), C.POINTER(TuneReport), vp]
        L.hpo_server_opts_default.argtypes = [C.POINTER(ServerOpts)]
        L.hpo_server_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(ServerOpts), C.POINTER(vp)]
        L.hpo_server_destroy.argtypes = [vp]
        L.hpo_server_open_stream.argtypes = [vp, C.c_int32, C.c_int64, _dp, C.c_int64, _dp, C.c_double,
                                             C.POINTER(C.c_int64)]
        L.hpo_server_close_stream.argtypes = [vp, C.c_int64]
        L.hpo_server_submit.argtypes = [vp, C.c_int64, C.c_double, _dp, C.c_int64, _ip, C.POINTER(C.c_int64)]
        L.hpo_server_wait.argtypes = [vp, C.c_int64, C.c_double, C.POINTER(RequestInfo)]
        L.hpo_server_stream_get.argtypes = [vp, C.c_int64, _dp, C.c_int64]
        L.hpo_server_hbm_used.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise HpoError(f"hpc_ode error {rc}: {lib().hpo_last_error().decode()}")


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


CELL_LAYOUTS = {0: "auto", 1: "wave", 2: "block"}


REACTORS = {"CONP": 0, "CONV": 1}


def _reactor(r):
    """0/1 or the reference's names: "CONP" constant pressure, "CONV" constant volume."""
    if isinstance(r, str):
        if r not in REACTORS:
            raise ValueError(f"reactor must be one of {sorted(REACTORS)}")
        return REACTORS[r]
    return int(r)


def opts(rtol=1e-6, atol=1e-10, maxsteps=10000, lu_precision=0, refine_sweeps=1,
         waves_per_block=0, blocks_per_cu=0, sort_cells=0, lu_layout=0, cell_layout=0, reactor=0,
         fp64_fallback=1) -> Opts:
    o = Opts()
    lib().hpo_opts_default(C.byref(o))
    o.rtol, o.atol, o.maxsteps = rtol, atol, maxsteps
    o.lu_precision, o.refine_sweeps = lu_precision, refine_sweeps
    o.waves_per_block, o.blocks_per_cu, o.sort_cells = waves_per_block, blocks_per_cu, sort_cells
    o.lu_layout = lu_layout
    o.cell_layout = cell_layout
    o.reactor = _reactor(reactor)
    o.fp64_fallback = fp64_fallback
    return o


class Mechanism:
    """A chemistry mechanism loaded from a Cantera YAML or CTML (.xml) file (the reference's
    resources/mechanisms/), with its tables resident on the GPU."""

    def __init__(self, path: str):
        h = C.c_void_p()
        _check(lib().hpo_mech_load(path.encode(), C.byref(h)))
        self._h = h
        self.path = path
        K, n, R = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().hpo_mech_info(h, C.byref(K), C.byref(n), C.byref(R)))
        self.n_species, self.n_state, self.n_reactions = K.value, n.value, R.value
        buf = C.create_string_buffer(64)
        self.species = []
        for k in range(self.n_species):
            _check(lib().hpo_mech_species_name(h, k, buf, 64))
            self.species.append(buf.value.decode())
        self.W = np.zeros(self.n_species)
        _check(lib().hpo_mech_molecular_weights(h, _p(self.W)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.hpo_mech_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def tables(self):
        K, R = self.n_species, self.n_reactions
        nasa = np.zeros((K, 14))
        tmid = np.zeros(K)
        arr = np.zeros((R, 8))
        kind = np.zeros(R, np.int32)
        rev = np.zeros(R, np.int32)
        _check(lib().hpo_mech_tables(self._h, _p(nasa), _p(tmid), _p(arr), _p(kind, _ip), _p(rev, _ip)))
        return dict(nasa=nasa, tmid=tmid, arr=arr, kind=kind, rev=rev)

    def enthalpy_mass(self, T, Y):
        """h(T, Y) in J/kg for temperatures T[n] and full mass-fraction rows Y[n, n_species]."""
        T = np.ascontiguousarray(np.atleast_1d(T), np.float64)
        Y = np.ascontiguousarray(np.atleast_2d(Y), np.float64)
        if Y.shape != (len(T), self.n_species):
            raise ValueError("Y must be [len(T), n_species]")
        h = np.zeros(len(T))
        _check(lib().hpo_mech_enthalpy_mass(self._h, len(T), _p(T), _p(Y), Y.shape[1], _p(h)))
        return h

    # --- component evaluations (GPU) -------------------------------------
    def rhs(self, y, P, cell_layout=0, reactor=0):
        """dydt on the GPU; cell_layout as in integrate (0 auto, 1 wave, 2 block); reactor "CONP"
        (P = pressure, Pa) or "CONV" (P = density, kg/m^3)."""
        y = np.ascontiguousarray(np.atleast_2d(y), np.float64)
        n = y.shape[0]
        P = np.ascontiguousarray(np.broadcast_to(np.asarray(P, np.float64), (n,)))
        f = np.zeros_like(y)
        r = _reactor(reactor)
        if cell_layout == 0 and r == 0:
            _check(lib().hpo_eval_rhs(self._h, n, _p(y), _p(P), _p(f)))
        else:
            _check(lib().hpo_eval_layout(self._h, cell_layout, r, 0, n, _p(y), _p(P), _p(f)))
        return f

    def jac(self, y, P, cell_layout=0, reactor=0):
        """d dydt / dy on the GPU, [n, n_state, n_state]; cell_layout and reactor as in rhs."""
        y = np.ascontiguousarray(np.atleast_2d(y), np.float64)
        n = y.shape[0]
        P = np.ascontiguousarray(np.broadcast_to(np.asarray(P, np.float64), (n,)))
        J = np.zeros((n, self.n_state, self.n_state))
        r = _reactor(reactor)
        if cell_layout == 0 and r == 0:
            _check(lib().hpo_eval_jac(self._h, n, _p(y), _p(P), _p(J)))
        else:
            _check(lib().hpo_eval_layout(self._h, cell_layout, r, 1, n, _p(y), _p(P), _p(J)))
        return J

    def density(self, y, P):
        """rho = P Wbar / (R T) of states y [n, n_state] at pressures P (kg/m^3): the value a
        constant-volume cell that starts at (T, Y, P) keeps (reactor "CONV" takes it as P)."""
        y = np.atleast_2d(np.asarray(y, np.float64))
        S = (y[:, 1:] / self.W[:-1]).sum(axis=1) + (1.0 - y[:, 1:].sum(axis=1)) / self.W[-1]
        return np.asarray(P, np.float64) / (8314.4621 * y[:, 0] * S)

    # --- integration ------------------------------------------------------
    def integrate(self, y, P, dt, rtol=1e-6, atol=1e-10, maxsteps=10000, lu_precision=0,
                  refine_sweeps=1, waves_per_block=0, blocks_per_cu=0, sort_cells=0,
                  return_timing=False, lu_layout=0, cell_layout=0, reactor=0, fp64_fallback=1):
        """Integrates every row of y ([ncells, n_state]) from t=0 to dt on the GPU. reactor
        "CONP": P is each cell's pressure; "CONV": P is each cell's density (see density()).
        Returns (y_out, status, stats[, timing])."""
        y = np.ascontiguousarray(np.atleast_2d(y), np.float64)
        n = y.shape[0]
        b = Batch(self, n)
        b.put(y, P, dt)
        t = b.integrate(n, opts(rtol, atol, maxsteps, lu_precision, refine_sweeps, waves_per_block,
                                blocks_per_cu, sort_cells, lu_layout, cell_layout, reactor, fp64_fallback))
        yo, st, stats = b.get(n)
        return (yo, st, stats, t) if return_timing else (yo, st, stats)


class Batch:
    """Device-resident cell states (hpo_batch)."""

    def __init__(self, mech: Mechanism, capacity: int):
        self.mech = mech
        h = C.c_void_p()
        _check(lib().hpo_batch_create(mech.handle, capacity, C.byref(h)))
        self._h = h
        self.capacity = capacity

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.hpo_batch_destroy(h)
            self._h = None

    def put(self, y, P, dt, first=0):
        y = np.ascontiguousarray(np.atleast_2d(y), np.float64)
        n = y.shape[0]
        Pa = np.ascontiguousarray(np.broadcast_to(np.asarray(P, np.float64), (n,)))
        da = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (n,)))
        _check(lib().hpo_batch_put(self._h, first, n, _p(y), y.shape[1], _p(Pa), 0.0, _p(da), 0.0))

    def integrate(self, count, o: Opts | None = None, stream=None) -> Timing:
        t = Timing()
        _check(lib().hpo_batch_integrate(self._h, count, C.byref(o) if o is not None else None,
                                         stream, C.byref(t)))
        return t

    def autotune(self, count, o: Opts, sample=32768, stream=None) -> TuneReport:
        """Tunes launch shape, queue order and (FP32 path) Newton-matrix layout for cells
        [0, count); o comes back tuned (sort_cells = 2 selects the tuned order in integrate)."""
        rep = TuneReport()
        _check(lib().hpo_batch_autotune(self._h, count, sample, C.byref(o), C.byref(rep), stream))
        return rep

    def get(self, count, first=0):
        n = self.mech.n_state
        y = np.zeros((count, n))
        st = np.zeros(count, np.int32)
        stats = np.zeros((count, NSTATS), np.int32)
        _check(lib().hpo_batch_get(self._h, first, count, _p(y), n, _p(st, _ip), _p(stats, _ip)))
        return y, st, {f: stats[:, i].copy() for i, f in enumerate(STATS_FIELDS)}


class Server:
    """Serving loop (hpo_server): resident cell streams, substep requests, continuous batching."""

    def __init__(self, mechs, n_slots=8, max_batch_cells=262144, hbm_budget_bytes=0, cu_share_pct=0, launch_cells=None,
                 **opt_kw):
        so = ServerOpts()
        lib().hpo_server_opts_default(C.byref(so))
        so.n_slots, so.max_batch_cells, so.hbm_budget_bytes = n_slots, max_batch_cells, hbm_budget_bytes
        so.cu_share_pct = cu_share_pct
        if launch_cells is not None:
            so.launch_cells = launch_cells
        opt_kw.setdefault("sort_cells", 1)   # the server's defaults: hottest cells first in each launch,
        opt_kw.setdefault("lu_precision", 1)  # FP32 Newton matrix with FP64 refinement and fallback
        so.opts = opts(**opt_kw)
        self.mechs = list(mechs)
        arr = (C.c_void_p * len(self.mechs))(*[m.handle.value for m in self.mechs])
        h = C.c_void_p()
        _check(lib().hpo_server_create(arr, len(self.mechs), C.byref(so), C.byref(h)))
        self._h = h
        self._pending = {}   # request id -> (y_out, status_out): keeps the buffers alive

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hpo_server_destroy(self._h)
            self._h = None

    __del__ = close

    def open_stream(self, mech_index, y0, P):
        """Returns a stream id, or None when the HBM budget is exhausted."""
        y0 = np.ascontiguousarray(np.atleast_2d(y0), np.float64)
        n = y0.shape[0]
        Pa = np.ascontiguousarray(np.broadcast_to(np.asarray(P, np.float64), (n,)))
        sid = C.c_int64()
        rc = lib().hpo_server_open_stream(self._h, mech_index, n, _p(y0), y0.shape[1], _p(Pa), 0.0, C.byref(sid))
        if rc == E_BUSY:
            return None
        _check(rc)
        return sid.value

    def close_stream(self, sid):
        _check(lib().hpo_server_close_stream(self._h, sid))

    def submit(self, sid, dt, want_state=True, n=None, n_state=None):
        """Enqueues one substep; returns a request id. With want_state the result rows and
        statuses are delivered to wait()."""
        y = st = None
        if want_state:
            y = np.zeros((n, n_state))
            st = np.zeros(n, np.int32)
        rid = C.c_int64()
        _check(lib().hpo_server_submit(self._h, sid, dt, _p(y) if y is not None else None,
                                       n_state if y is not None else 0, _p(st, _ip) if st is not None else None,
                                       C.byref(rid)))
        self._pending[rid.value] = (y, st)
        return rid.value

    def wait(self, rid, timeout=-1.0):
        info = RequestInfo()
        _check(lib().hpo_server_wait(self._h, rid, timeout, C.byref(info)))
        y, st = self._pending.pop(rid)
        return y, st, info.as_dict()

    def stream_state(self, sid, n, n_state):
        y = np.zeros((n, n_state))
        _check(lib().hpo_server_stream_get(self._h, sid, _p(y), n_state))
        return y

    def hbm(self):
        u, b = C.c_int64(), C.c_int64()
        _check(lib().hpo_server_hbm_used(self._h, C.byref(u), C.byref(b)))
        return u.value, b.value


def lu_solve(M, b, fp32=False, layout=None, timing=False):
    """Batched LU + solve (parity tests). layout None: hpo_lu_solve, the BDF kernel's register
    code; 0 / 1: hpo_lu_solve_layout with one wave / one 256-thread workgroup per matrix;
    2: the block-per-cell kernel's Gauss-Jordan code (ns 53 or 100).
    timing=True also returns the launch's device time in ms (layouts 0 and 1)."""
    M = np.ascontiguousarray(M, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n, ns, _ = M.shape
    x = np.zeros((n, ns))
    ok = np.zeros(n, np.int32)
    if layout is None:
        _check(lib().hpo_lu_solve(ns, int(fp32), n, _p(M), _p(b), _p(x), _p(ok, _ip)))
        return x, ok
    ms = C.c_float()
    _check(lib().hpo_lu_solve_layout(ns, int(fp32), int(layout), n, _p(M), _p(b), _p(x), _p(ok, _ip),
                                     C.byref(ms)))
    return (x, ok, ms.value) if timing else (x, ok)


def device_info():
    cu = C.c_int32()
    fr, to = C.c_int64(), C.c_int64()
    _check(lib().hpo_device_info(C.byref(cu), C.byref(fr), C.byref(to)))
    return dict(n_cu=cu.value, free_bytes=fr.value, total_bytes=to.value)


def exported_symbols():
    """Names every ABI function declared in include/hpc_ode.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "hpc_ode.h")
    with open(hdr) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(hpo_[a-z_0-9]+)\s*\(", txt)))
