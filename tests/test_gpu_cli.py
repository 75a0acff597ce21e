"""The glemur executable under the reference's run-directory contract
(src/main.cu:44-394, src/io.cu): params.in and data/ in the working
directory; cuda.info, data/B0.vtk, data/dump<k>.vtk, data/save.vtk and
data/time_series.dat written; fRestart = t resumes from save.vtk."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "glemur_amd", "glemur")

PARAMS = """&comp
    nx = 17, ny = 19, nz = 23
/
&start
    Lx = 4, Ly = 4.5, Lz = 12, Ox = -2, Oy = -2.25, Oz = -6
    bInit = "blobs", ampl = 1, nBlobs = 6
    blobXc = 0.5, -0.5, 0.5, -0.5, 0.5, -0.5
    blobYc = 6*0
    blobZc = -5, -3, -1, 1, 3, 5
    blobScale = 6*0.8, blobZl = 6*1
    blobTwist = 0.4, -0.4, 0.4, -0.4, 0.4, -0.4
    fRestart = {restart}
/
&run
    nCycle = {ncycle}, maxError = 1e-5, dt0 = 3e-4, dtMin = 1e-9
    epsilonProf = f, pressure = f, zUpdate = t, jMethod = "Stokes"{run_extra}
/
&io
    nTs = 2, dtDump = 0.05, dtSave = 0.1
    dumpJ = t, dumpDetJac = t, dumpCellVol = t, dumpConvexity = t, dumpWedgeMin = t
    redB2 = t, redJMax = t, redJxB_B2Max = t, redConvex = t, redWedgeMin = t, redEpsilonStar = t
    redErrB_1ez = t{io_extra}
/
"""


def run_cli(d, *args, restart=False, ncycle=30, run_extra="", io_extra="", ok=True):
    (d / "params.in").write_text(PARAMS.format(restart="t" if restart else "f", ncycle=ncycle, run_extra=run_extra,
                                               io_extra=io_extra))
    os.makedirs(d / "data", exist_ok=True)
    print("running glemur", args, "in", d, flush=True)
    r = subprocess.run([EXE, *args], cwd=d, capture_output=True, text=True, timeout=120)
    if ok:
        assert r.returncode == 0, r.stdout + r.stderr
    else:
        assert r.returncode != 0, r.stdout
    return r.stdout + r.stderr


def ts_rows(d):
    lines = (d / "data" / "time_series.dat").read_text().splitlines()
    assert lines[0].startswith("#")
    return lines[0].split(), [l.split() for l in lines[1:]]


@pytest.mark.parametrize("prec", ["64", "32"])
def test_run_directory_contract(require_gpu, tmp_path, prec):
    out = run_cli(tmp_path, "-d", "0", "-p", prec)
    assert "glemur terminated successfully" in out
    assert (tmp_path / "cuda.info").stat().st_size > 0
    assert (tmp_path / "data" / "B0.vtk").exists()
    assert (tmp_path / "data" / "save.vtk").exists()
    dumps = [f for f in os.listdir(tmp_path / "data") if f.startswith("dump")]
    assert dumps
    raw = (tmp_path / "data" / dumps[0]).read_bytes()
    for name in (b"SCALARS cellVol", b"SCALARS convexity", b"SCALARS wedgeMin"):
        assert name in raw
    head, rows = ts_rows(tmp_path)
    assert head[:5] == ["#", "it", "t", "dt", "maxDelta"] or head[:4] == ["#it", "t", "dt", "maxDelta"]
    # nTs = 2: a row on every second accepted step (time index 0, 2, ...)
    its = [int(r[0]) for r in rows]
    assert its == list(range(0, 30, 2))
    t = [float(r[1]) for r in rows]
    assert all(b > a for a, b in zip(t, t[1:]))
    col = lambda name: [float(r[head.index(name) - (1 if head[0] == "#" else 0)]) for r in rows]
    b2 = col("B2")
    assert all(b <= a for a, b in zip(b2, b2[1:]))
    assert all(v == 1.0 for v in col("convex")) and all(v > 0 for v in col("wedgeMin"))
    assert all(v > 0 for v in col("epsilonStar")) and all(v > 0 for v in col("errB_1ez"))


def test_restart_continues(require_gpu, tmp_path):
    """A run of 30 steps, then fRestart = t for 10 more: the restarted run
    starts from save.vtk's time and keeps the time series going."""
    run_cli(tmp_path, "-p", "64", ncycle=30)
    import glemur_amd as g
    p = g.params_read(tmp_path / "params.in")
    xb, B0, t_save, dt_save = g.read_state(tmp_path, p, 64)
    assert t_save > 0 and dt_save > 0
    _, rows_before = ts_rows(tmp_path)
    run_cli(tmp_path, "-p", "64", restart=True, ncycle=10)
    _, rows_after = ts_rows(tmp_path)
    assert len(rows_after) > len(rows_before)
    t_new = [float(r[1]) for r in rows_after[len(rows_before):]]
    assert min(t_new) > t_save
    B0_init, _ = g.init_state(p)
    assert np.array_equal(B0, B0_init)


INERTIA = dict(run_extra=", inertia = t, nu = 0.5", io_extra=", redU2 = t, redUMax = t")


def test_inertia_run_and_restart(require_gpu, tmp_path):
    """inertia = t through the executable: the velocity starts at rest and
    is driven by J x B (U2 and UMax grow from zero) and save.vtk holds a
    ufield block.  The save keeps the whole FP64 state (positions with their
    ghosts, velocity, t, dt), so 30 steps plus a restart for 10 more print
    the same time-series rows as 40 steps in one run.  FP32 runs refuse
    inertia."""
    one, two = tmp_path / "one", tmp_path / "two"
    for d in (one, two):
        os.makedirs(d)
    run_cli(one, "-p", "64", ncycle=40, **INERTIA)
    run_cli(two, "-p", "64", ncycle=30, **INERTIA)
    head, rows = ts_rows(two)
    col = lambda name, rs: [float(r[head.index(name) - (1 if head[0] == "#" else 0)]) for r in rs]
    u2, umax = col("U2", rows), col("UMax", rows)
    assert u2[-1] > 0 and umax[-1] > 0 and max(u2) > u2[0]
    assert b"VECTORS ufield" in (two / "data" / "save.vtk").read_bytes()
    import glemur_amd as g
    p = g.params_read(two / "params.in")
    xb, B0, t_save, dt_save, uu = g.read_state(two, p, 64)
    assert np.abs(uu).max() > 0
    run_cli(two, "-p", "64", restart=True, ncycle=10, **INERTIA)
    _, rows_two = ts_rows(two)
    _, rows_one = ts_rows(one)
    cont = rows_two[len(rows):]
    want = [r for r in rows_one if int(r[0]) >= 30]
    assert len(cont) == len(want) == 5
    # the time index restarts at 0; every other column is printed identically
    assert [r[1:] for r in cont] == [r[1:] for r in want], (cont, want)
    out = run_cli(tmp_path, "-p", "32", ncycle=5, ok=False, **INERTIA)
    assert "inertia" in out


NO_EXTRA = ", redConvex = f, redWedgeMin = f, redEpsilonStar = f, redErrB_1ez = f"


@pytest.mark.parametrize("prec", ["64", "32"])
def test_batched_rows_from_step_log(require_gpu, tmp_path, prec):
    """Without a column that only the separate reductions give, the
    executable runs the attempts in device batches (stopping only for dumps,
    saves and the end) and takes each row's B2, JMax and JxB_B2Max from the
    step log: the diagnostics recorded at the start of the attempt after an
    accepted step.  With such a column it stops every nTs accepted steps and
    reduces the fields.  Both runs take the same steps: the same rows (time
    index, t, dt, maxDelta printed identically; the diagnostics, printed to
    six digits, within one unit of the last: the stage kernels reduce in
    another order, and the FP32 ones in FP32, what the fields pass reduces in
    FP64) and the same dumps and save, byte for byte."""
    a, b = tmp_path / "log", tmp_path / "red"
    for d in (a, b):
        os.makedirs(d)
    out = run_cli(a, "-p", prec, ncycle=40, io_extra=NO_EXTRA)
    assert "time stepping: " in out
    run_cli(b, "-p", prec, ncycle=40)
    ha, ra = ts_rows(a)
    hb, rb = ts_rows(b)
    assert len(ra) == len(rb) == 20
    off = lambda h: 1 if h[0] == "#" else 0
    for x, y in zip(ra, rb):
        assert x[:4] == y[:4], (x, y)
        for name in ("B2", "JMax", "JxB_B2Max"):
            va, vb = float(x[ha.index(name) - off(ha)]), float(y[hb.index(name) - off(hb)])
            assert abs(va - vb) <= 1.01e-5 * abs(vb), (name, va, vb)
    da = sorted(f for f in os.listdir(a / "data") if f.endswith(".vtk"))
    db = sorted(f for f in os.listdir(b / "data") if f.endswith(".vtk"))
    assert da == db and "save.vtk" in da and any(f.startswith("dump") for f in da)
    for f in da:
        assert (a / "data" / f).read_bytes() == (b / "data" / f).read_bytes(), f


@pytest.mark.parametrize("faces", ["line-tied", "periodic"])
def test_both_precisions_through_the_executable(require_gpu, tmp_path, faces):
    """The executable with all faces line-tied and with xPeri = zPeri = t
    (both precisions step on the fused stage kernels; FP32 with -p 32).
    Both precisions write every row.  From dt0 =
    3e-4 the adaptive dt grows at a rate set by the error estimate, which
    FP32 rounding dominates while dt is small, so the two runs take their own
    step sequences and reach each row at their own t; compared at the same t
    (linear interpolation between FP64 rows), their energies B2 differ by the
    time discretisation of two different adaptive sequences (each step's
    local error up to maxError), under 2e-4 relative here, alike for both
    kinds of faces; the FP32 energy never rises."""
    extra = ", xPeri = t, zPeri = t" if faces == "periodic" else ""
    rows = {}
    for prec in ("64", "32"):
        d = tmp_path / prec
        os.makedirs(d)
        run_cli(d, "-p", prec, ncycle=30, run_extra=extra)
        h, r = ts_rows(d)
        off = 1 if h[0] == "#" else 0
        rows[prec] = np.array([(float(x[h.index("t") - off]), float(x[h.index("B2") - off])) for x in r])
        assert (d / "data" / "save.vtk").exists()
    a, b = rows["64"], rows["32"]
    assert len(a) == len(b) == 15
    assert np.all(np.diff(b[:, 1]) <= 0)
    inside = (b[:, 0] >= a[0, 0]) & (b[:, 0] <= a[-1, 0])
    assert inside.sum() >= 5
    want = np.interp(b[inside, 0], a[:, 0], a[:, 1])
    err = np.max(np.abs(b[inside, 1] - want) / want)
    print(f"{faces}: B2 of the FP32 run against the FP64 run at the same t: {err:.2e} ({inside.sum()} rows); "
          f"t at the last row {a[-1, 0]:.4f} (FP64), {b[-1, 0]:.4f} (FP32)")
    assert err < 2e-4
