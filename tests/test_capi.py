"""The C-ABI library loads and exports every symbol include/*.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(couplb_\w+)\s*\(", txt)
    return sorted(set(n for n in names if not n.endswith("_eval")))


def test_headers_declare_entry_points():
    syms = declared_symbols()
    assert "couplb_lbm_step" in syms and "couplb_fix_post_force" in syms
    assert "couplb_ibm_mdf_device" in syms and "couplb_mlp_train_device" in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol():
    import couplb_amd
    lib = ctypes.CDLL(couplb_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_string():
    import couplb_amd
    assert "gfx950" in couplb_amd.version()


def test_python_constants_match_the_header():
    # the binding's status, storage, kernel, step-flag and MDF method codes
    # are the values include/couplb_hip.h defines
    import couplb_amd
    txt = open(os.path.join(ROOT, "include", "couplb_hip.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define COUPLB_(\w+)\s+(-?\d+)", txt)}
    pairs = {"OK": "OK", "EINVAL": "EINVAL", "EDEVICE": "EDEVICE", "ENOMEM": "ENOMEM", "ESTATE": "ESTATE",
             "FP64": "FP64", "FP32": "FP32", "FP16": "FP16", "BF16": "BF16", "I16": "I16",
             "ROMA": "ROMA", "PESKIN4": "PESKIN4", "STEP_RHO_REFRESHED_NEXT": "RHO_REFRESHED_NEXT",
             "MDF_RICHARDSON": "MDF_RICHARDSON", "MDF_CG": "MDF_CG", "MDF_PCG": "MDF_PCG"}
    for h, p in pairs.items():
        assert defs[h] == getattr(couplb_amd, p), h
