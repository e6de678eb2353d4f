"""Host-side checks (no GPU): the C-ABI library loads and exports exactly
what include/clm_soil.h declares; argument validation fails cleanly before
any device call; the ctypes binding mirrors the header; the synthetic
workload generator is deterministic, prefix-stable and matches its spec."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "clm_soil.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(clm_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from clmsoil import _lib
    lib = C.CDLL(_lib.LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 18
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == decl
    out = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    exported = sorted(set(re.findall(r" T (clm_\w+)", out)))
    assert exported == decl


def test_header_enums_match_binding():
    from clmsoil import _lib
    txt = open(HEADER).read()
    assert "CLM_ABI_VERSION 3" in txt
    lib = _lib.load()
    assert lib.clm_abi_version() == 3
    assert "CLM_NPARAM_RRE 6" in txt and _lib.NPARAM_RRE == 6 and _lib.NPARAM_EH == 13
    block = txt.split("CLM_P_NU = 0,")[1].split("};")[0]
    order = re.findall(r"CLM_P_(\w+)", block)
    assert ["NU"] + order == ["NU", "THETA_R", "VG_ALPHA", "VG_N", "KSAT", "SS", "NU_SS_OM",
                                   "NU_SS_QUARTZ", "NU_SS_GRAVEL", "KAPPA_DRY",
                                   "KAPPA_SAT_FROZEN", "KAPPA_SAT_UNFROZEN", "RHOC_DS"]


def test_struct_layouts_match_header(tmp_path):
    """Size and every field offset of the ctypes mirrors against what a C
    compiler makes of include/clm_soil.h."""
    import subprocess
    from clmsoil import _lib
    structs = {"clm_config": _lib.Config, "clm_state": _lib.State, "clm_aux": _lib.Aux,
               "clm_surface_fluxes": _lib.SurfaceFluxes, "clm_step_stats": _lib.StepStats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "clm_soil.h"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                             text=True).stdout.splitlines():
        cname, field, val = ln.split()
        got[(cname, field)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


@pytest.mark.parametrize("mutate,code", [
    (lambda c: setattr(c, "model", 7), -1),
    (lambda c: setattr(c, "bc_bottom", 5), -1),
    (lambda c: setattr(c, "nparam", 3), -1),
    (lambda c: setattr(c, "jac_precision", 16), -1),
    (lambda c: setattr(c, "newton_max_iters", 0), -1),
    (lambda c: setattr(c, "solver", 9), -1),
    (lambda c: setattr(c, "jacobian", 5), -1),
    (lambda c: setattr(c, "jacobian", 1), -1),        # the Newton matrix is EH only
    (lambda c: setattr(c, "bc_top", 1), -1),          # free drainage is a bottom condition
    (lambda c: setattr(c, "bc_top", 7), -1),
    (lambda c: (setattr(c, "model", 1), setattr(c, "nparam", 13), setattr(c, "bc_top", 2)), -1),
    (lambda c: setattr(c, "bc_top", 3), -1),          # TOPMODEL needs 9 parameter rows
    (lambda c: (setattr(c, "model", 1), setattr(c, "nparam", 13), setattr(c, "bc_top", 3)), -1),
    (lambda c: (setattr(c, "model", 1), setattr(c, "nparam", 13), setattr(c, "bc_top", 4)), -1),
    (lambda c: setattr(c, "bc_top", 5), -1),          # AtmosDrivenFluxBC is EH only
    # EH + TOPMODEL atmos needs 16 parameter rows
    (lambda c: (setattr(c, "model", 1), setattr(c, "nparam", 13), setattr(c, "bc_top", 6)), -1),
    # EH + atmos on the lane solver
    (lambda c: (setattr(c, "model", 1), setattr(c, "nparam", 13), setattr(c, "bc_top", 5),
                setattr(c, "solver", 2)), -1),
])
def test_create_rejects_bad_config_without_device(mutate, code):
    """Validation runs before any HIP call, so this is safe without a GPU."""
    from clmsoil import _lib
    lib = _lib.load()
    cp = np.array([0, 3, 7], dtype=np.int32)
    zf = np.array([-1, -0.5, -0.2, 0, -2, -1.5, -1, -0.5, 0.0])
    P = np.ones((6, 2))
    cfg = _lib.Config(0, 1, 2, _lib.iptr(cp), _lib.dptr(zf), _lib.dptr(P), 6, 64, 3, 0.0, 0)
    mutate(cfg)
    ctx = C.c_void_p()
    assert lib.clm_create(C.byref(ctx), C.byref(cfg), None) == code
    assert not ctx
    assert lib.clm_last_error(None)


def test_create_rejects_bad_geometry_without_device():
    from clmsoil import _lib
    lib = _lib.load()
    cp = np.array([0, 3, 4], dtype=np.int32)   # second column has 1 layer
    zf = np.linspace(-1, 0, 6)
    P = np.ones((6, 2))
    cfg = _lib.Config(0, 1, 2, _lib.iptr(cp), _lib.dptr(zf), _lib.dptr(P), 6, 64, 3, 0.0, 0)
    ctx = C.c_void_p()
    assert lib.clm_create(C.byref(ctx), C.byref(cfg), None) == -1
    assert b"layers" in lib.clm_last_error(None)
    cp = np.array([0, 3], dtype=np.int32)
    zf = np.array([0.0, -0.5, -0.7, -1.0])      # decreasing faces
    cfg = _lib.Config(0, 1, 1, _lib.iptr(cp), _lib.dptr(zf), _lib.dptr(np.ones((6, 1))), 6, 64,
                      3, 0.0, 0)
    assert lib.clm_create(C.byref(ctx), C.byref(cfg), None) == -1
    assert b"increasing" in lib.clm_last_error(None)


def test_generator_is_deterministic_and_prefix_stable():
    from clmsoil import inputs
    a = inputs.make_batch("rre_131072_uneven", ncol=500)
    b = inputs.make_batch("rre_131072_uneven", ncol=500)
    c = inputs.make_batch("rre_131072_uneven", ncol=200)
    assert np.array_equal(a.vartheta_l, b.vartheta_l) and np.array_equal(a.params, b.params)
    assert np.array_equal(a.params[:, :200], c.params)
    assert np.array_equal(a.col_ptr[:201], c.col_ptr)
    assert np.array_equal(a.z_face[:c.z_face.size], c.z_face)
    assert np.array_equal(a.top_water[:, :200], c.top_water)
    sub = a.column_subset(np.arange(200))
    assert np.array_equal(sub.vartheta_l, c.vartheta_l)


def test_splitmix64_known_values():
    """splitmix64 (Steele, Lea, Flood 2014) reference outputs for seed 0."""
    from clmsoil.inputs import ColumnRNG
    r = ColumnRNG(0, 1)
    out = [int(r.next_u64()[0]) for _ in range(3)]
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


@pytest.mark.parametrize("name", ["rre_4096x50", "rre_131072_uneven", "eh_65536x20",
                                  "eh_262144_uneven"])
def test_workload_matches_spec(name):
    from clmsoil import inputs
    b = inputs.make_batch(name, ncol=3000)
    w = b.workload
    P = b.params
    assert P[0].min() >= 0.35 and P[0].max() <= 0.55
    assert P[1].min() >= 0.02 and P[1].max() <= 0.10
    assert P[2].min() >= 0.5 and P[2].max() <= 10.0
    assert P[3].min() >= 1.2 and P[3].max() <= 3.0
    assert P[4].min() >= 1e-7 and P[4].max() <= 1e-4
    assert np.all(P[5] == 1e-3)
    nl = b.nlayers
    assert nl.min() >= w.layers[0] and nl.max() <= w.layers[1]
    if w.layers[0] != w.layers[1]:
        assert len(np.unique(nl)) > 0.9 * (w.layers[1] - w.layers[0] + 1)
    for c in range(0, 3000, 499):
        zf = b.z_face[b.col_ptr[c] + c:b.col_ptr[c + 1] + c + 1]
        assert zf[-1] == 0.0 and np.all(np.diff(zf) > 0)
        D = -zf[0]
        assert w.depth[0] - 1e-12 <= D <= w.depth[1] + 1e-12
        dz = np.diff(zf)[::-1]  # from the top down
        if w.stretch is None:
            assert np.allclose(dz, D / nl[c])
        else:
            assert np.allclose(dz[1:] / dz[:-1], w.stretch, rtol=1e-9)
    s = (b.vartheta_l - np.repeat(P[1], nl)) / np.repeat(P[0] - P[1], nl)
    assert s.min() >= 0.3 - 1e-12 and s.max() <= 0.9 + 1e-12
    assert b.nslots == -(-w.nsteps // w.steps_per_slot)
    if w.top_flux is not None:
        assert b.top_water.min() >= w.top_flux[0] and b.top_water.max() <= w.top_flux[1]
    if b.model == 1:
        assert np.all(b.theta_i == 0)
        assert P[12].min() >= 1.5e6 and P[12].max() <= 2.5e6
        assert P[9].min() >= 0.2 and P[9].max() <= 0.4
        h = b.top_heat.ravel()
        assert abs(h.mean()) < 5.0 and 70 < h.std() < 90
