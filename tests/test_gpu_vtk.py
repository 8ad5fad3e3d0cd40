"""VTK output of the MI355X fix (`vtk`, `vtk_region`, `vtk_solid`) against
the oracle's FixCoupLB on the same deck: the fluid fields of every
prefix_STEP.vti equal the oracle grid's rho, u (compute_macroscopic(false))
and force fields Grid::fx/fy/fz at end_of_step, scaled as write_vtk does
(couplb_io.h:42-275) -- including the unfolded periodic images in the middle
of an LB interval and the folded ones after its last sub-step
(fix_couplb.cpp:795-797) -- and the .vtp / .pvd files list what
write_solid_vtk / write_pvd_impl list (couplb_io.h:304-647)."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle
from tests.scenarios import atoms_from, parse_cmd

pytestmark = pytest.mark.gpu

N = (12, 10, 8)
BASE = "f all couplb 12 10 8 0.1 1.0 md_per_lb 2 kernel peskin4 xi_ibm 0.5 gravity 1e-5 0 0"


def markers():
    rng = np.random.default_rng(7)
    # near the periodic x and z faces (ghost images matter) and in the bulk
    x = np.array([[0.3, 4.2, 3.7], [11.8, 5.1, 4.4], [6.1, 3.3, 0.2], [5.5, 6.6, 7.9],
                  [0.1, 2.5, 0.4], [8.2, 8.7, 5.1], [3.3, 1.2, 2.2], [9.9, 4.4, 6.6]])
    v = rng.normal(0, 1e-3, x.shape)
    return atoms_from(x, v, np.ones(len(x)))


def read_vti(path):
    root = ET.parse(path).getroot()
    img = root.find("ImageData")
    arrays = {}
    for da in img.find("Piece").find("CellData").findall("DataArray"):
        vals = np.array(da.text.split(), dtype=np.int64 if da.get("type") == "Int32" else np.float64)
        nc = int(da.get("NumberOfComponents", "1"))
        arrays[da.get("Name")] = vals.reshape(-1, nc) if nc > 1 else vals
    return img.attrib, img.find("Piece").attrib, arrays


def oracle_fields(fo):
    o, g = fo.ops_ptr.contents, fo.g
    o.compute_macroscopic(g, 0)
    gx, gy, gz = fo.gdims
    sl = (slice(1, gz - 1), slice(1, gy - 1), slice(1, gx - 1))
    f = lambda k: fo.field(k).reshape(gz, gy, gx)[sl]
    return (f(oracle.F_RHO), np.stack([f(oracle.F_UX), f(oracle.F_UY), f(oracle.F_UZ)], -1),
            np.stack([f(oracle.F_FX), f(oracle.F_FY), f(oracle.F_FZ)], -1))


def close(a, b, what):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    print(f"  {what}: max |diff| / max = {err:.2e}", flush=True)
    assert err < 2e-8, what   # the files hold 9 significant digits (%.8e)


def test_vtk_fields_solid_and_series_match_the_oracle(tmp_path):
    import couplb_amd
    prefix = str(tmp_path / "flow")
    cmd = (BASE + f" vtk 1 {prefix} vtk_region 0 12 2 9 0 8 "
           "vtk_solid id type x y z vx fx")
    print("case VTK: 12x10x8, md_per_lb 2, 8 markers near periodic faces", flush=True)
    a_dev, a_ora = markers(), markers()
    a_dev["type"] = np.ones(len(a_dev["mass"]), np.int32)
    fo = oracle.Fix(dim=3, dt=0.5, **parse_cmd(BASE))
    fd = couplb_amd.Fix(cmd, dt=0.5)
    fd.init()
    tags = np.arange(1, len(a_dev["mass"]) + 1, dtype=np.int32)
    fd.set_atom_extras(tag=tags)
    fo.setup(a_ora)
    fd.setup(a_dev)
    vel_scale, force_scale = fd.scalar("vel_scale"), fd.scalar("force_scale")
    steps = [0]
    for step in range(1, 5):
        for a in (a_dev, a_ora):
            a["f"][:] = 0.0
        fo.post_force(a_ora)
        fd.post_force(a_dev)
        fo.end_of_step(step)
        fd.end_of_step(step)
        steps.append(step)
        path = f"{prefix}_{step:06d}.vti"
        assert os.path.exists(path)
        attrs, piece, arr = read_vti(path)
        assert attrs["WholeExtent"] == piece["Extent"] == "0 12 0 7 0 8"
        org = [float(t) for t in attrs["Origin"].split()]
        assert org == [0.0, 2.0, 0.0] and attrs["Spacing"].split() == ["1.00000000e+00"] * 3
        rho, u, F = oracle_fields(fo)
        reg = (slice(0, 8), slice(2, 9), slice(0, 12))
        mid = step % 2 == 1   # after the first of two sub-steps: images not folded yet
        print(f"step {step} ({'mid-interval' if mid else 'interval end'})", flush=True)
        close(arr["rho"].reshape(8, 7, 12), rho[reg] * 1.0, "rho")
        close(arr["velocity_phys"].reshape(8, 7, 12, 3), u[reg] * vel_scale, "velocity_phys")
        Fv = arr["force_phys"].reshape(8, 7, 12, 3)
        close(Fv, F[reg] * force_scale, "force_phys")
        assert np.abs(F[reg]).max() > 0
        assert (arr["type"] == 0).all()
        # the atoms of the group inside the snapped region [0,12) x [2,9) x [0,8)
        root = ET.parse(f"{prefix}_solid_{step:06d}.vtp").getroot()
        pc = root.find("PolyData").find("Piece")
        x = a_dev["x"]
        inside = (x[:, 1] >= 2) & (x[:, 1] < 9)
        assert int(pc.get("NumberOfPoints")) == inside.sum()
        pts = np.array(pc.find("Points").find("DataArray").text.split(), float).reshape(-1, 3)
        assert np.allclose(pts, x[inside], rtol=1e-8, atol=0)
        pd = {d.get("Name"): (d.get("NumberOfComponents"),
                              np.array(d.text.split(), float)) for d in pc.find("PointData")}
        assert list(pd) == ["id", "type", "position", "vx", "fx"]
        assert pd["position"][0] == "3" and pd["id"][0] is None
        assert np.array_equal(pd["id"][1], tags[inside])
        assert np.allclose(pd["vx"][1], a_dev["v"][inside, 0], rtol=1e-8, atol=1e-20)
        assert np.allclose(pd["fx"][1], a_dev["f"][inside, 0], rtol=1e-8, atol=1e-20)
        assert np.allclose(pd["fx"][1], a_ora["f"][inside, 0], rtol=1e-7, atol=1e-14)
    # collections: every written step, timestep = step * dt, names without directory
    for pvd, ext in ((prefix + ".pvd", "vti"), (prefix + "_solid.pvd", "vtp")):
        ds = ET.parse(pvd).getroot().find("Collection").findall("DataSet")
        stem = os.path.basename(prefix) + ("_solid" if ext == "vtp" else "")
        assert [d.get("file") for d in ds] == [f"{stem}_{s:06d}.{ext}" for s in steps]
        assert [float(d.get("timestep")) for d in ds] == [0.5 * s for s in steps]
    fd.close()
    fo.close()


def test_node_forces_fold_only_moves_ghost_images():
    # device export: folded minus unfolded is exactly the ghost-image part,
    # which vanishes away from the periodic faces
    import couplb_amd
    d = couplb_amd.Device(3, *N, tau=0.8, kernel=couplb_amd.PESKIN4)
    d.init_equilibrium(1.0, (0.01, 0.0, 0.0))
    a = markers()
    f = np.zeros_like(a["x"])
    d.ibm(a["x"], a["v"], a["mass"], f, (0, 0, 0), 1.0, 1.0, 1.0, 1.0, 1.0)
    Fu, Ff = d.node_forces(fold=False), d.node_forces(fold=True)
    diff = np.abs(Ff - Fu).max(-1)            # [z, y, x]
    assert diff.max() > 0
    inner = diff[2:-2, 2:-2, 2:-2]            # two or more cells from every (periodic) face
    assert inner.max() == 0
    # spreading conserves the total force only once the images are folded in
    assert np.allclose(Ff.reshape(-1, 3).sum(0), -f.sum(0), rtol=1e-12, atol=1e-18)
    assert not np.allclose(Fu.reshape(-1, 3).sum(0), -f.sum(0), rtol=1e-6, atol=0)
