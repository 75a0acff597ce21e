"""Output files of the run-directory contract (src/io.cu:167-386, 390-499,
503-678): byte layout checked with a small legacy-VTK parser written here
from the format (text header lines, big-endian binary blocks), restart
round trip, and the time-series columns.  CPU only."""
import os

import numpy as np
import pytest

import glemur_amd as g
from glemur_amd import workloads


def small_params(**kw):
    p = workloads.braid_params(6, 5, 7)
    for k, v in kw.items():
        setattr(p, k, v)
    return g.params_derive(p)


class Reader:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def line(self):
        e = self.b.index(b"\n", self.o)
        s = self.b[self.o:e].decode()
        self.o = e + 1
        return s

    def block(self, dtype, n):
        a = np.frombuffer(self.b, dtype=dtype, count=n, offset=self.o)
        self.o += a.nbytes
        return a

    def done(self):
        return self.o == len(self.b)


def expect_param_block(r, p, t, dt, fdt, rs):
    assert r.line() == "FIELD parameters 9"
    assert r.line() == f"t 1 1 {rs}"
    assert r.block(fdt, 1)[0] == t
    assert r.line() == f"dt 1 1 {rs}"
    assert r.block(fdt, 1)[0] == dt
    assert r.line() == "nx_ny_nz 1 3 int"
    assert list(r.block(">i4", 3)) == [p.nx, p.ny, p.nz]
    for name, vals in (("Lx_Ly_Lz", (p.Lx, p.Ly, p.Lz)), ("Ox_Oy_Oz", (p.Ox, p.Oy, p.Oz)),
                       ("dx_dy_dz", (p.dx, p.dy, p.dz)), ("ampl", (p.ampl,)), ("phi1_phi2", (p.phi1, p.phi2)),
                       ("maxError", (p.maxError,))):
        assert r.line() == f"{name} 1 {len(vals)} {rs}"
        np.testing.assert_array_equal(r.block(fdt, len(vals)), np.array(vals, dtype=fdt))


@pytest.mark.parametrize("bits", [64, 32])
def test_b0_vtk_layout(tmp_path, bits):
    p = small_params()
    os.makedirs(tmp_path / "data")
    B0, xb = g.init_state(p)
    g.write_b0(tmp_path, p, B0, bits)
    fdt, rs = (">f8", "double") if bits == 64 else (">f4", "float")
    r = Reader(tmp_path / "data" / "B0.vtk")
    assert r.line() == "# vtk DataFile Version 2.0"
    assert r.line() == "GLEMuR B0 dump"
    assert r.line() == "BINARY"
    assert r.line() == "DATASET STRUCTURED_POINTS"
    assert r.line() == "DIMENSIONS %9i %9i %9i" % (p.nx + 2, p.ny + 2, p.nz + 2)
    assert r.line() == "ORIGIN %8.12f %8.12f %8.12f" % (p.Ox - p.dx, p.Oy - p.dy, p.Oz - p.dz)
    assert r.line() == "SPACING %8.12f %8.12f %8.12f" % (p.Lx / (p.nx - 1), p.Ly / (p.ny - 1), p.Lz / (p.nz - 1))
    assert r.line() == "POINT_DATA %9i" % p.npad
    assert r.line() == f"VECTORS bfield {rs}"
    np.testing.assert_array_equal(r.block(fdt, 3 * p.npad), B0.astype(fdt))
    assert r.done()


@pytest.mark.parametrize("bits", [64, 32])
def test_dump_vtk_layout(tmp_path, bits):
    p = small_params(dumpJ=True, dumpDetJac=True, dumpCellVol=True, dumpConvexity=True, dumpWedgeMin=True)
    os.makedirs(tmp_path / "data")
    rng = np.random.default_rng(0)
    xb = rng.standard_normal(3 * p.npad)
    B = rng.standard_normal(3 * p.npad)
    J = rng.standard_normal(3 * p.ninner)
    det = rng.random(p.npad)
    mesh = {k: rng.standard_normal(p.ninner) for k in ("cellVol", "convexity", "wedgeMin")}
    t, dt = 1.25, 3.5e-3
    g.dump_state(tmp_path, p, t, dt, 4, xb, B, J, det, bits, **mesh)
    fdt, rs = (">f8", "double") if bits == 64 else (">f4", "float")
    r = Reader(tmp_path / "data" / "dump4.vtk")
    assert r.line() == "# vtk DataFile Version 2.0"
    assert r.line() == "GLEMuR data dump"
    assert r.line() == "BINARY"
    assert r.line() == "DATASET STRUCTURED_GRID"
    assert r.line() == "DIMENSIONS %9i %9i %9i" % (p.nx + 2, p.ny + 2, p.nz + 2)
    expect_param_block(r, p, t, dt, fdt, rs)
    assert r.line() == "POINTS %9i %s" % (p.npad, rs)
    np.testing.assert_array_equal(r.block(fdt, 3 * p.npad), xb.astype(fdt))
    assert r.line() == "POINT_DATA %9i" % p.npad
    assert r.line() == f"VECTORS bfield {rs}"
    np.testing.assert_array_equal(r.block(fdt, 3 * p.npad), B.astype(fdt))
    assert r.line() == f"VECTORS jfield {rs}"
    jf = r.block(fdt, 3 * p.npad).reshape(p.nz + 2, p.ny + 2, p.nx + 2, 3)
    want = np.zeros_like(jf)
    want[1:-1, 1:-1, 1:-1] = J.reshape(p.nz, p.ny, p.nx, 3).astype(fdt)
    np.testing.assert_array_equal(jf, want)
    assert r.line() == f"SCALARS detJac {rs}"
    assert r.line() == "LOOKUP_TABLE default"
    np.testing.assert_array_equal(r.block(fdt, p.npad), det.astype(fdt))
    for name in ("cellVol", "convexity", "wedgeMin"):     # src/io.cu:334-385: ghosts are zeros
        assert r.line() == f"SCALARS {name} {rs}"
        assert r.line() == "LOOKUP_TABLE default"
        v = r.block(fdt, p.npad).reshape(p.nz + 2, p.ny + 2, p.nx + 2)
        want = np.zeros_like(v)
        want[1:-1, 1:-1, 1:-1] = mesh[name].reshape(p.nz, p.ny, p.nx).astype(fdt)
        np.testing.assert_array_equal(v, want)
    assert r.done()


def test_save_and_restart_round_trip(tmp_path):
    """dumpState(n=-1) + writeB0 then readState gives back xb, B0, t, dt
    bitwise (src/io.cu:390-499)."""
    p = small_params()
    os.makedirs(tmp_path / "data")
    B0, xb = g.init_state(p)
    xb = xb + 1e-3 * np.random.default_rng(1).standard_normal(xb.shape)
    g.write_b0(tmp_path, p, B0, 64)
    g.dump_state(tmp_path, p, 0.75, 1.5e-4, -1, xb, B0)
    assert os.path.exists(tmp_path / "data" / "save.vtk")
    xr, Br, t, dt = g.read_state(tmp_path, p, 64)
    assert np.array_equal(xr, xb) and np.array_equal(Br, B0)
    assert t == 0.75 and dt == 1.5e-4


def test_velocity_dump_and_restart(tmp_path):
    """inertia = t: the velocity is written as "VECTORS ufield" between
    bfield and jfield with zero ghosts (src/io.cu:269-290) and read back on
    restart (src/io.cu:466-479) bitwise; no ufield without inertia."""
    p = small_params(inertia=True, dumpJ=True)
    os.makedirs(tmp_path / "data")
    rng = np.random.default_rng(4)
    B0, xb = g.init_state(p)
    B = rng.standard_normal(3 * p.npad)
    uu = rng.standard_normal(3 * p.ninner)
    J = rng.standard_normal(3 * p.ninner)
    g.write_b0(tmp_path, p, B0, 64)
    g.dump_state(tmp_path, p, 0.5, 2e-3, -1, xb, B, J, uu=uu)
    r = Reader(tmp_path / "data" / "save.vtk")
    for _ in range(5):
        r.line()
    expect_param_block(r, p, 0.5, 2e-3, ">f8", "double")
    assert r.line() == "POINTS %9i double" % p.npad
    r.block(">f8", 3 * p.npad)
    assert r.line() == "POINT_DATA %9i" % p.npad
    assert r.line() == "VECTORS bfield double"
    np.testing.assert_array_equal(r.block(">f8", 3 * p.npad), B)
    assert r.line() == "VECTORS ufield double"
    uf = r.block(">f8", 3 * p.npad).reshape(p.nz + 2, p.ny + 2, p.nx + 2, 3)
    want = np.zeros_like(uf)
    want[1:-1, 1:-1, 1:-1] = uu.reshape(p.nz, p.ny, p.nx, 3)
    np.testing.assert_array_equal(uf, want)
    assert r.line() == "VECTORS jfield double"
    xr, Br, t, dt, ur = g.read_state(tmp_path, p, 64)
    assert np.array_equal(xr, xb) and np.array_equal(Br, B0) and np.array_equal(ur, uu)
    assert t == 0.5 and dt == 2e-3
    # a save without the velocity cannot restart an inertia run
    q = small_params(dumpJ=True)
    g.dump_state(tmp_path, q, 0.5, 2e-3, -1, xb, B, J, uu=uu)
    assert b"ufield" not in open(tmp_path / "data" / "save.vtk", "rb").read()
    with pytest.raises(g.GlemurError):
        g.read_state(tmp_path, p, 64)


def test_read_state_missing_file_is_an_error(tmp_path):
    p = small_params()
    with pytest.raises(g.GlemurError):
        g.read_state(tmp_path, p, 64)


def test_time_series_columns(tmp_path):
    """Header written when t == 0, one row per call after; columns in the
    reference order, plus the two force-freeness columns this library adds
    last."""
    p = small_params()
    p.redJMax = p.redJxB_B2Max = p.redB2 = True
    p.redEpsilonStar = p.redConvex = False
    os.makedirs(tmp_path / "data")
    g.write_ts(tmp_path, p, None, 0.0, 0.0, 0, 0.0)
    red = g.Reduction()
    red.JMax, red.JxB_B2Max, red.B2, red.JxB_JBMax, red.JxBMax_JBMax = 2.5, 0.125, 384.0, 0.0625, 1.5e-3
    g.write_ts(tmp_path, p, red, 0.5, 1e-3, 7, 3.25e-6)
    lines = open(tmp_path / "data" / "time_series.dat").read().splitlines()
    assert lines[0] == "#%7s%13s%13s%13s%13s%13s%13s%13s%13s" % ("it", "t", "dt", "maxDelta", "JMax", "JxB_B2Max",
                                                                  "B2", "JxB_JBMax", "JxBMax_JBMax")
    assert lines[1] == "%8i %12.5e %12.5e %12.5e %12.5e %12.5e %12.5e %12.5e %12.5e " % (
        7, 0.5, 1e-3, 3.25e-6, 2.5, 0.125, 384.0, 0.0625, 1.5e-3)
    # without the extra column the row is the reference's
    g.write_ts(tmp_path, p, None, 0.0, 0.0, 0, 0.0, jxb_jb_col=False)
    g.write_ts(tmp_path, p, red, 0.5, 1e-3, 7, 3.25e-6, jxb_jb_col=False)
    lines = open(tmp_path / "data" / "time_series.dat").read().splitlines()
    assert lines[0].split() == ["#", "it", "t", "dt", "maxDelta", "JMax", "JxB_B2Max", "B2"] or \
        lines[0].split() == ["#it", "t", "dt", "maxDelta", "JMax", "JxB_B2Max", "B2"]
    assert len(lines[1].split()) == 7


def test_init_matches_oracle_bitwise():
    """Host initial state of the library equals the oracle's restatement of
    src/init.cu bit for bit, for every closed-form field it supports."""
    import oracle as o
    for name, kw in (("blobs", {}), ("Pontin09", {}), ("analytic", {}), ("homZ", {}),
                     ("blobs", {"initDist": "initShearX", "initShearA": 0.05, "initShearB": 0.02}),
                     ("blobs", {"initDist": "centerShift", "initShearA": 0.05})):
        p = small_params()
        p.set_str("bInit", name)
        for k, v in kw.items():
            if isinstance(v, str):
                p.set_str(k, v)
            else:
                setattr(p, k, v)
        g.params_derive(p)
        B0, xb = g.init_state(p)
        oB0, oxb = o.init_state(o.from_product(p))
        assert np.array_equal(B0, oB0), name
        assert np.array_equal(xb, oxb), name
