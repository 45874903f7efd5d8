"""The C-ABI library loads and exports every function the public headers declare; argument
validation rejects inconsistent meshes before touching the device (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from pihm_amd import _lib
from pihm_amd.meshgen import make_watershed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    pat = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\(", re.M)
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"^\s*#.*$", "", txt, flags=re.M)
        txt = re.sub(r"struct\s+_generic_N_Vector_Ops\s*\{.*?\};", "", txt, flags=re.S)
        for m in pat.finditer(txt):
            n = m.group(1)
            if n in ("typedef", "return", "sizeof", "if", "defined", "int", "void", "double", "long"):
                continue
            names.add(n)
    names.discard("PihmRhsFn")
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    declared = _declared()
    assert len(declared) > 80
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    for n in declared:
        getattr(L, n)
    # the bindings used by the Python host side are all declared
    assert set(_lib.exported_symbols()) <= set(declared)


def _ctx(desc):
    ctx = C.c_void_p()
    return _lib.lib().pihm_ctx_create(C.byref(desc), C.byref(ctx)), ctx


def test_ctx_rejects_bad_meshes():
    m = make_watershed(535, seed=1, n_seg=20)
    d = m.desc()
    d.nelem = 0
    assert _ctx(d)[0] == _lib.PIHM_ERR_ARG
    d = m.desc()
    d.surf_mode = 3
    assert _ctx(d)[0] == _lib.PIHM_ERR_ARG
    d = m.desc()
    d.stepsize = 0.0
    assert _ctx(d)[0] == _lib.PIHM_ERR_ARG
    # unknown channel shape (the reference exits at run time, src/river_flow.c:541-544)
    bad = m.a["intrpl_ord"].copy()
    bad[3] = 7
    d = m.desc()
    d.intrpl_ord = bad.ctypes.data_as(_lib.c_int32_p)
    assert _ctx(d)[0] == _lib.PIHM_ERR_ARG
    # bank element without the river tag on any edge
    left = m.a["leftele"].copy()
    left[0] = 1 if left[0] != 1 else 2
    d = m.desc()
    d.leftele = left.ctypes.data_as(_lib.c_int32_p)
    assert _ctx(d)[0] == _lib.PIHM_ERR_ARG
    # unknown outlet condition
    down = m.a["down"].copy()
    down[down < 0] = -9
    d = m.desc()
    d.down = down.ctypes.data_as(_lib.c_int32_p)
    assert _ctx(d)[0] == _lib.PIHM_ERR_ARG
    # neighbour index out of range
    nabr = m.a["nabr"].copy()
    nabr[0, 0] = m.nelem + 5
    d = m.desc()
    d.nabr = nabr.ctypes.data_as(_lib.c_int32_p)
    assert _ctx(d)[0] == _lib.PIHM_ERR_ARG


def test_integrator_rejects_bad_calls():
    L = _lib.lib()
    assert not L.PihmCVodeCreate(1, 2)          # only BDF
    mem = L.PihmCVodeCreate(2, 2)
    assert mem
    assert L.PihmCVodeSStolerances(mem, -1.0, 1e-4) == -22
    assert L.PihmCVSpgmr(mem, 1, 5) == -22      # preconditioning is not offered (PREC_NONE)
    p = C.c_void_p(mem)
    L.PihmCVodeFree(C.byref(p))
    assert not p.value
