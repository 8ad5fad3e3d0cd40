"""Host-side FixCoupLB core (include/couplb_fix.h) without a GPU: argument
parsing and the init-time checks that fire before any device work, with the
reference's messages (fix_couplb.cpp:71-238, 256-349)."""
import numpy as np
import pytest

import couplb_amd
from couplb_amd import CoupLBError, Fix


@pytest.mark.parametrize("cmd,msg", [
    ("f all couplb 8 8", "need Nx Ny Nz nu rho"),
    ("f all couplb 0 8 8 0.1 1.0", "must be positive"),
    ("f all couplb 8 8 8 -0.1 1.0", "viscosity must be positive"),
    ("f all couplb 8 8 8 0.1 0.0", "density must be positive"),
    ("f all couplb 8 8 8 0.1 1.0 bogus 1", "Unknown fix couplb keyword: bogus"),
    ("f all couplb 8 8 8 0.1 1.0 md_per_lb 0", "md_per_lb must be >= 1"),
    ("f all couplb 8 8 8 0.1 1.0 xi_ibm 5.0", "0 < xi_ibm <= 1"),
    ("f all couplb 8 8 8 0.1 1.0 wall_y 1 5", "types must be in [0,4]"),
    ("f all couplb 8 8 8 0.1 1.0 kernel gauss", "use roma or peskin4"),
    ("f all couplb 8 8 8 0.1 1.0 gravity 1 2", "gravity: need 3 values"),
    ("f all couplb 8 8 8 0.1 1.0 check_every -1", "check_every cannot be negative"),
    ("f all couplb 8 8 8 0.1 1.0 vtk_region 0 1 2", "vtk_region: need 4 values"),
    ("f all couplb 8 8 8 0.1 1.0 storage fp8", "use fp64, fp32, fp16, bf16 or i16"),
        ("f all couplb 8 8 8 0.1 1.0 storage i16 storage_range 0", "storage_range must be > 0"),
    ("f all couplb 8 8 8 0.1 1.0 vtk 10 v vtk_solid id vtk_solid x", "vtk_solid: keyword specified more than once"),
    ("f all couplb 8 8 8 0.1 1.0 vtk_solid", "vtk_solid: need at least one attribute"),
    ("f all couplb 8 8 8 0.1 1.0 vtk -1 v", "vtk frequency cannot be negative"),
])
def test_constructor_errors(cmd, msg):
    with pytest.raises(CoupLBError, match=msg.replace("[", r"\[").replace("]", r"\]")):
        Fix(cmd)


def test_2d_requires_nz_1():
    with pytest.raises(CoupLBError, match="Nz must be 1 for 2D"):
        Fix("f all couplb 8 8 4 0.1 1.0", dimension=2)


@pytest.mark.parametrize("cmd,kw,msg", [
    ("f all couplb 8 8 8 0.1 1.0", dict(boxhi=(8, 9, 8)), "non-cubic grid cells"),
    ("f all couplb 8 8 8 1e-20 1.0", {}, "tau <= 0.5"),
    ("f all couplb 8 8 8 0.1 1.0 wall_vel 0.4 0 0", {}, "wall Ma > 0.5"),
    ("f all couplb 8 32 8 0.1 1.0 wall_y 1 1 gravity 1e-3 0 0", {}, "Poiseuille Ma > 0.5"),
    ("f all couplb 8 8 8 0.1 1.0 gravity v_gx 0 0", {}, "Variable gx for gravity x"),
    # vtk_solid / vtk_region set-up, fix_couplb.cpp:378-425
    ("f all couplb 8 8 8 0.1 1.0 vtk_solid id", {}, "vtk_solid requires vtk keyword"),
    ("f all couplb 8 8 8 0.1 1.0 vtk 10 v vtk_region 5 3 0 8 0 8", {}, "empty or inverted region"),
    # clamped to the box first: a region beyond the box is empty
    ("f all couplb 8 8 8 0.1 1.0 vtk 10 v vtk_region 9 12 0 8 0 8", {},
     r"empty or inverted region \(\[8,8\]x\[0,8\]x\[0,8\]\)"),
])
def test_init_errors_before_device(cmd, kw, msg):
    fx = Fix(cmd, **kw)
    with pytest.raises(CoupLBError, match=msg):
        fx.init()


def test_setmask_is_post_force_and_end_of_step():
    assert couplb_amd.lib().couplb_fix_setmask() == 3


def test_valid_deck_parses_all_keywords():
    fx = Fix("flow all couplb 16 64 1 0.1 1.0 wall_y 1 1 gravity 5e-6 0.0 0.0 kernel roma "
             "output 400 prof.dat check_every 200 md_per_lb 2 xi_ibm 0.5 wall_x 0 0 "
             "wall_vel 0.0 0.0 0.0 vtk 100 v vtk_region 0 1 0 1 vtk_solid id "
             "checkpoint 1000 ck restart ck storage fp32 wall_vel_face y hi 0.01 0 0",
             dimension=2)
    assert fx.scalar("tau") == 0.0  # not initialised yet
