"""The C-ABI library loads on a machine without a GPU and exports every function
that include/*.h declares (no compute calls here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "hydrocode-fusion_amd", "lib")
# which shared library implements each header
HEADER_LIB = {"hydrocode_grp.h": "libhydrogrp.so", "hydrocode_dropin.h": "libhydrocode2d.so"}

DECL = re.compile(r"^\s*(?:extern\s+)?(?:const\s+)?(?:int|void|double|_Bool|hgrp_bool|struct\s+\w+)\s*\*?\s*(\w+)\s*\(",
                  re.MULTILINE)


def declared_functions():
    names = {}
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.DOTALL)
        src = re.sub(r"typedef[^;]*;", "", src, flags=re.DOTALL)
        for name in DECL.findall(src):
            names[name] = os.path.basename(h)
    return names


def test_headers_declare_the_boundary():
    names = declared_functions()
    for required in ("GRP_solver_2D_split_EUL_source", "GRP_solver_2D_EUL_source", "GRP_solver_EUL_source",
                     "GRP_2D_flux", "linear_GRP_solver_Edir_Q1D", "linear_GRP_solver_Edir", "Riemann_solver_exact",
                     "bound_cond_slope_limiter_x", "bound_cond_slope_limiter_y", "flux_generator_x",
                     "flux_generator_y", "hgrp2d_create", "hgrp2d_solve", "hgrp2d_bench", "hgrp_grp2d_flux_batch",
                     "hgrp2d_slopes", "hgrp2d_fluxes", "hgrp1d_run_until"):
        assert required in names, required


def test_library_exports_every_declared_symbol():
    libs = {}
    for so in set(HEADER_LIB.values()):
        path = os.path.join(LIBDIR, so)
        if not os.path.exists(path):
            pytest.fail(f"{path} not built: run __graft_entry__.build()")
        libs[so] = ctypes.CDLL(path)
    names = declared_functions()
    assert set(names.values()) <= set(HEADER_LIB) | {"var_struc.h"}, set(names.values())
    missing = [(n, h) for n, h in names.items() if not hasattr(libs[HEADER_LIB[h]], n)]
    assert not missing, missing
