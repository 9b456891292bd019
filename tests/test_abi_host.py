"""Host-side checks of the C ABI (no GPU calls): the library loads, exports every function
include/hpc_ode.h declares, parses mechanisms like the oracle does, and reports errors."""
import os
import ctypes as C
import subprocess

import numpy as np
import pytest

import hpc_ode_amd as H
import oracle as O


def test_library_exports_every_declared_symbol():
    syms = H.exported_symbols()
    assert len(syms) >= 18
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing


def test_no_oracle_in_product_library():
    out = subprocess.run(["nm", "-D", H.LIB_PATH], capture_output=True, text=True).stdout
    assert "or_cvode" not in out and "or_rhs" not in out
    ldd = subprocess.run(["ldd", H.LIB_PATH], capture_output=True, text=True).stdout
    assert "liboracle" not in ldd and "pyjac" not in ldd


@pytest.mark.parametrize("which", ["gri", "h2", "h2li", "est"])
def test_cpp_reader_matches_oracle_reader(which, gri_path, h2_path, h2li_path):
    """The product's C++ readers (YAML; CTML for h2_li_2004.xml and ESTiMatE-Mech, whose PLOG
    reactions report their lowest-pressure level here) and the oracle's readers (PyYAML;
    xml.etree) give identical tables."""
    est = os.path.join(os.path.dirname(h2_path), "ESTiMatE-Mech_04042022.xml")
    path = {"gri": gri_path, "h2": h2_path, "h2li": h2li_path, "est": est}[which]
    m = H.Mechanism(path)
    o = O.Oracle(path)
    assert m.species == o.mech.species
    assert np.array_equal(m.W, o.t.W)
    t = m.tables()
    assert np.array_equal(t["tmid"], o.t.tmid)
    assert np.array_equal(t["nasa"], o.t.nasa)
    assert np.array_equal(t["kind"], o.t.kind)
    assert np.array_equal(t["rev"], o.t.rev)
    fall = o.t.kind >= 2
    # lnA, b, theta, A of every reaction; low-pressure parameters of falloff reactions
    assert np.allclose(t["arr"][:, :4], o.t.arr[:, :4], rtol=1e-15, atol=0)
    assert np.allclose(t["arr"][fall, 4:], o.t.arr[fall, 4:], rtol=1e-15, atol=0)


def test_reader_errors_are_reported(tmp_path):
    with pytest.raises(H.HpoError, match="cannot open"):
        H.Mechanism(str(tmp_path / "missing.yaml"))
    bad = tmp_path / "bad.yaml"
    bad.write_text("phases:\n- name: x\n  species: [A]\nspecies:\n- name: A\n  composition: {Zz: 1}\n"
                   "  thermo: {model: NASA7, temperature-ranges: [200, 1000, 3000], data: [[1,2,3,4,5,6,7]]}\n"
                   "reactions: []\n")
    with pytest.raises(H.HpoError, match="unknown element"):
        H.Mechanism(str(bad))


def test_invalid_arguments_return_error_codes():
    L = H.lib()
    assert L.hpo_mech_load_yaml(None, None) == -1
    assert L.hpo_batch_create(None, 10, None) == -1
    o = H.Opts()
    L.hpo_opts_default(C.byref(o))
    assert o.rtol == 1e-6 and o.atol == 1e-10 and o.maxsteps == 10000
    assert L.hpo_lu_solve(7, 0, 1, None, None, None, None) == -1
    # the layout checks come before any device call (HPO_E_INVALID = -1, HPO_E_UNSUPPORTED = -5)
    M, b, x = np.eye(10)[None], np.ones((1, 10)), np.zeros((1, 10))
    ok = np.zeros(1, np.int32)
    args = [M.ctypes.data_as(C.POINTER(C.c_double)), b.ctypes.data_as(C.POINTER(C.c_double)),
            x.ctypes.data_as(C.POINTER(C.c_double)), ok.ctypes.data_as(C.POINTER(C.c_int32)), None]
    assert L.hpo_lu_solve_layout(10, 0, 1, 1, None, None, None, None, None) == -1
    assert L.hpo_lu_solve_layout(10, 0, 3, 1, *args) == -1
    assert L.hpo_lu_solve_layout(10, 1, 1, 1, *args) == -5
    assert L.hpo_lu_solve_layout(7, 0, 1, 1, *args) == -5
    assert L.hpo_lu_solve_layout(10, 0, 2, 1, *args) == -5   # the block code exists for 53 and 100


def test_eval_layout_checks_come_before_device_calls(h2_path):
    m = H.Mechanism(h2_path)
    L = H.lib()
    y, P, out = np.zeros((1, m.n_state)), np.ones(1), np.zeros((1, m.n_state))
    dp = C.POINTER(C.c_double)
    a = [y.ctypes.data_as(dp), P.ctypes.data_as(dp), out.ctypes.data_as(dp)]
    assert L.hpo_eval_layout(m.handle, 3, 0, 0, 1, *a) == -1    # cell layout
    assert L.hpo_eval_layout(m.handle, 1, 2, 0, 1, *a) == -1    # reactor
    assert L.hpo_eval_layout(m.handle, 1, 0, 2, 1, *a) == -1    # what
    assert L.hpo_eval_layout(m.handle, 2, 0, 0, 1, *a) == -5    # no block kernel for 10 variables


def test_reactor_option_checks(h2_path):
    """hpo_opts::reactor outside 0..1 is refused before any device work; the Python names map to
    the reference's CONP / CONV."""
    assert H.opts(reactor="CONV").reactor == 1 and H.opts(reactor="CONP").reactor == 0
    with pytest.raises(ValueError):
        H.opts(reactor="CONSTANT")
    m = H.Mechanism(h2_path)
    o = H.opts(reactor=2)
    fake = C.c_void_p(4096)   # never dereferenced: the option check comes first
    rc = H.lib().hpo_integrate_device(m.handle, 1, fake, fake, m.n_state, fake, fake, fake, None, None, None,
                                      C.byref(o), None, None)
    assert rc == -1 and b"reactor" in H.lib().hpo_last_error()


def test_server_option_checks_come_before_device_calls(h2_path):
    """hpo_server_create rejects a CU share outside 0..100 and too many slots before it touches
    the device."""
    m = H.Mechanism(h2_path)
    L = H.lib()
    arr = (C.c_void_p * 1)(m.handle.value)
    for field, bad in (("cu_share_pct", 101), ("cu_share_pct", -5), ("n_slots", 17)):
        so = H.ServerOpts()
        L.hpo_server_opts_default(C.byref(so))
        assert so.cu_share_pct == 0 and so.n_slots == 8 and so.opts.sort_cells == 1
        assert so.opts.lu_precision == 1 and so.opts.fp64_fallback == 1
        setattr(so, field, bad)
        h = C.c_void_p()
        assert L.hpo_server_create(arr, 1, C.byref(so), C.byref(h)) == -1
        assert not h.value
        assert field.encode() in L.hpo_last_error()


STRUCTS = [("hpo_opts", "Opts"), ("hpo_timing", "Timing"), ("hpo_tune_report", "TuneReport"),
           ("hpo_server_opts", "ServerOpts"), ("hpo_request_info", "RequestInfo")]


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every struct of include/hpc_ode.h has the same size and field offsets in its ctypes
    mirror: a C program compiled against the header prints them."""
    import os
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "hpc_ode.h"', "int main(void) {"]
    for cname, pyname in STRUCTS:
        lines.append(f'  printf("{pyname} size %zu\\n", sizeof({cname}));')
        for field, _ in getattr(H, pyname)._fields_:
            lines.append(f'  printf("{pyname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for _, pyname in STRUCTS:
        cls = getattr(H, pyname)
        assert got[(pyname, "size")] == C.sizeof(cls), pyname
        for field, _ in cls._fields_:
            assert got[(pyname, field)] == getattr(cls, field).offset, (pyname, field)
