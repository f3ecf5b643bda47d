"""The reference's file formats (grid.dat, bcs.dat, control.dat, PETSc
binary Vec) on the CPU.  The PETSc Vec layout is pinned against bytes built
by hand from the published format (class id 1211214, big-endian 32-bit
header, big-endian doubles), not against the writer itself."""
import struct

import numpy as np
import pytest

from curvib_old_amd import cases, formats


def test_petsc_vec_bytes(tmp_path):
    p = tmp_path / "v.dat"
    formats.write_petsc_vec(str(p), np.array([1.0, -2.5, 3.0e-300]))
    assert p.read_bytes() == struct.pack(">ii3d", 1211214, 3, 1.0, -2.5, 3.0e-300)
    assert not (tmp_path / "v.dat.info").exists()


def test_petsc_vec_roundtrip_dof3(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((4, 5, 6, 3))
    p = str(tmp_path / "vfield00001_0.dat")
    formats.write_petsc_vec(p, a, block_size=3)
    assert open(p + ".info").read().strip() == "-vecload_block_size 3"
    b = formats.read_petsc_vec(p)
    assert np.array_equal(b, a.ravel())          # interleaved Cmpnts in natural order


def test_petsc_vec_rejects_other_files(tmp_path):
    p = tmp_path / "x.dat"
    p.write_bytes(struct.pack(">ii", 1211216, 3))      # a Mat class id
    with pytest.raises(ValueError):
        formats.read_petsc_vec(str(p))
    p.write_bytes(struct.pack(">ii2d", 1211214, 3, 1.0, 2.0))
    with pytest.raises(ValueError):
        formats.read_petsc_vec(str(p))


def test_grid_dat_roundtrip(tmp_path):
    case = cases.ChannelCase(cells=(6, 5, 4), amp=0.1)
    xyz = cases.node_coords(case)
    p = str(tmp_path / "grid.dat")
    formats.write_grid_dat(p, [xyz])
    lines = open(p).read().split("\n")
    assert lines[0] == "1" and lines[1] == "7 6 5"
    (IM, JM, KM, back), = formats.read_grid_dat(p)
    assert (IM, JM, KM) == case.nodes
    assert np.array_equal(back, xyz)             # %.17e round-trips doubles


def test_grid_dat_1d_layout(tmp_path):
    p = tmp_path / "grid.dat"
    x, y, z = [0.0, 0.5, 1.0], [0.0, 0.25, 0.75, 1.0], [0.0, 2.0]
    rows = ["1", "3 4 2"] + [f"{v} 9 9" for v in x] + [f"9 {v} 9" for v in y] + [f"9 9 {v}" for v in z]
    p.write_text("\n".join(rows) + "\n")
    (IM, JM, KM, xyz), = formats.read_grid_dat(str(p), grid1d=True, scale=2.0)
    assert (IM, JM, KM) == (3, 4, 2)
    assert xyz.shape == (2, 4, 3, 3)
    assert xyz[1, 2, 1].tolist() == [1.0, 1.5, 4.0]


def test_bcs_and_control(tmp_path):
    formats.write_bcs_dat(str(tmp_path / "bcs.dat"), [(7, 7, 1, 1, 7, 7)])
    assert formats.read_bcs_dat(str(tmp_path / "bcs.dat")) == [(7, 7, 1, 1, 7, 7)]
    (tmp_path / "control.dat").write_text(
        "-dt 0.06 #//0.0005\n-ren 778\n#-rstart 590\n\n-snes_monitor\n-mg_k_semi 1\n-mg_j_semi 0\n"
        "-ps_ksp_atol 1.e-5\n-ps_ksp_rtol 1.e-11\n-ps_ksp_max_it 40\n-x_c     0.\n-red_vel -0.5\n")
    o = formats.read_control_dat(str(tmp_path / "control.dat"))
    assert o["dt"] == "0.06" and o["snes_monitor"] is True and "rstart" not in o and o["red_vel"] == "-0.5"
    r = formats.run_options(o)
    assert r["ren"] == 778.0 and r["dt"] == 0.06 and r["semi"] == (0, 0, 1) and r["mglevels"] == 3
    s = formats.solver_options(o)
    assert s == {"rtol": 1e-11, "atol": 1e-5, "max_it": 40, "restart": 30}
    assert formats.solver_options({})["rtol"] == 1e-5          # PETSc KSP default


def test_icem_surface_round_trip(tmp_path):
    """nlist/elist (ibm_read_Icem, ibm_io.c:3076-3240) written and read back:
    vertices to the last bit, triangles, unit outward normals of a sphere;
    -char_length_ibm, -x_c.. and L_dim transform the vertices as read."""
    from curvib_old_amd import ibsearch
    v, t = ibsearch.icosphere((0.5, 0.4, 0.3), 0.2, 2)
    formats.write_icem(str(tmp_path), v, t, ibi=1)
    v2, t2, nf = formats.read_icem(str(tmp_path), ibi=1)
    assert np.array_equal(v, v2) and np.array_equal(t, t2)
    np.testing.assert_allclose(np.linalg.norm(nf, axis=1), 1.0, rtol=1e-14)
    out = v[t].mean(axis=1) - np.array([0.5, 0.4, 0.3])
    assert (np.einsum("nc,nc->n", nf, out) > 0).all()
    v3, _, _ = formats.read_icem(str(tmp_path), ibi=1, cl=2.0, center=(1.0, 0.0, 0.0), L_dim=3.0)
    np.testing.assert_allclose(v3, (v / 2.0 + np.array([1.0, 0.0, 0.0])) * 3.0, rtol=1e-15)
