"""The lsim driver on the GPU: the reference's command line and time loop over
the C ABI.  Its checkpoints and VTK output must equal, bit for bit, the same
steps run through the Python binding, and a restart from a checkpoint must
continue as the reference's does (src/main.cpp:151-181: n = restart_step,
t = the checkpoint's step time, so that step runs again)."""
import ctypes as C
import re
import struct
import subprocess
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import pytest

from leachingsim_amd import lsgpu as L
from tests.helpers import small_mixed_geometry

pytestmark = pytest.mark.gpu

PKG = Path(__file__).resolve().parents[1] / "leachingsim_amd"
LSIM = PKG / "lsim"
SHAPE = (24, 20, 28)      # (i, j, k) = (zext, yext, xext)
DT = 60.0


def run_lsim(tmp, raw, *extra):
    cmd = [str(LSIM), "--raw-file", str(raw), "--zext", str(SHAPE[0]), "--yext", str(SHAPE[1]),
           "--xext", str(SHAPE[2]), "--out_dir", str(tmp), "--dt", str(DT), "--bc", "zfaces",
           "--solver", "plain", "--warm", "previous", *extra]
    print("running", " ".join(cmd), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return out.stdout


def load_own_checkpoint(path):
    raw = Path(path).read_bytes()
    hdr = struct.calcsize("<8s6iqdd")
    magic, ver, ni, nj, nk, lb, _, step, t, dt = struct.unpack_from("<8s6iqdd", raw)
    assert magic == b"LSCHKPT1" and (ni, nj, nk) == SHAPE
    n = ni * nj * nk
    o = hdr
    img = np.frombuffer(raw, np.uint8, n, o); o += n * lb
    frac = np.frombuffer(raw, np.float32, n, o); o += 4 * n
    f = []
    for _ in range(5):
        f.append(np.frombuffer(raw, np.float64, n, o)); o += 8 * n
    return dict(step=step, t=t, img=img, frac=frac, inventory=f[0], saturation=f[1],
                pressure=f[2], conc=f[3], acid=f[4])


def python_run(img, times, restart_times=()):
    """Steps at `times` on one context; then, like a restart from a
    checkpoint, the state moves to a fresh context (no solver history: the
    transport's extrapolated start begins again from the old field) for the
    steps at `restart_times`."""
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_PLAIN64,
                         warm_start=L.WARM_PREVIOUS)
    g = L.GpuSimulation(SHAPE, P)
    g.read_labels(img)
    g.init_state()
    for t in times:
        assert g.step(t, DT).status == 0
    s = g.state()
    g.close()
    if restart_times:
        h = L.GpuSimulation(SHAPE, P)
        h.read_labels(img)
        h.init_state()
        for f, a in s.items():
            h.upload(f, a)
        for t in restart_times:
            assert h.step(t, DT).status == 0
        s = h.state()
        h.close()
    return s


def test_lsim_matches_binding_and_restarts(tmp_path):
    img = small_mixed_geometry(SHAPE, 31)
    raw = tmp_path / "geom.raw"
    img.astype(np.uint8).tofile(raw)
    log = run_lsim(tmp_path, raw, "--tmax", "240", "--nout", "2", "--chk_freq", "2")
    assert "*** Simulation Finished ***" in log and log.count("Done iteration.") == 4
    # the reference's per-solve lines, read with its benchmark parser's regex
    # (plot_benchmarks.py:22-23): one [Saturation] and one [Leaching] line per
    # step, whose sub-step counts it sums
    assert log.count("    [Pressure] Converged in ") == 4
    assert log.count("    [Concentration] Converged in ") == 4
    assert log.count("    [Acid Transport] Converged in ") == 4
    subs = [int(m) for m in re.findall(r"Sub-steps:\s+(\d+)", log)]
    assert len(subs) == 8 and all(n >= 1 for n in subs)
    ck = load_own_checkpoint(tmp_path / "checkpoint_000002.lsck")
    assert ck["step"] == 2 and ck["t"] == 2 * DT
    ref = python_run(img, [0.0, 60.0, 120.0])           # steps n = 0, 1, 2
    for key, f in (("img", L.F_IMG), ("frac", L.F_FRAC), ("inventory", L.F_INVENTORY),
                   ("saturation", L.F_SATURATION), ("pressure", L.F_PRESSURE),
                   ("conc", L.F_CONC), ("acid", L.F_ACID)):
        assert np.array_equal(ck[key].reshape(SHAPE), ref[f].reshape(SHAPE)), key
    # VTK output of step 2: pressure in VTK order (i fastest)
    m = ET.parse(tmp_path / "output_000002.pvti").getroot()
    assert m.find("PImageData").find("Piece").get("Source") == "output_0_000002.vti"
    raw_vti = (tmp_path / "output_0_000002.vti").read_bytes()
    head = raw_vti[:raw_vti.index(b"<AppendedData")].decode() + "</VTKFile>"
    das = {d.get("Name"): d for d in ET.fromstring(head).iter("DataArray")}
    data = raw_vti[raw_vti.index(b"_", raw_vti.index(b"<AppendedData")) + 1:]
    off = int(das["Pressure"].get("offset"))
    (nb,) = struct.unpack_from("<Q", data, off)
    p = np.frombuffer(data, np.float64, nb // 8, off + 8)
    assert np.array_equal(p, ref[L.F_PRESSURE].transpose(2, 1, 0).reshape(-1))
    pvd = ET.parse(tmp_path / "time_mapping.pvd").getroot()
    assert [d.get("file") for d in pvd.iter("DataSet")] == ["output_000000.pvti",
                                                          "output_000002.pvti"]
    # restart from step 2: step 2 runs again at t = 120 and its checkpoint is rewritten
    log = run_lsim(tmp_path, raw, "--tmax", "240", "--chk_freq", "2", "--nout", "100",
                   "--restart_step", "2", "--max_steps", "1")
    assert "Restarting simulation from step 2" in log
    ck2 = load_own_checkpoint(tmp_path / "checkpoint_000002.lsck")
    ref2 = python_run(img, [0.0, 60.0, 120.0], [120.0])
    for key, f in (("img", L.F_IMG), ("saturation", L.F_SATURATION),
                   ("pressure", L.F_PRESSURE), ("conc", L.F_CONC)):
        assert np.array_equal(ck2[key].reshape(SHAPE), ref2[f].reshape(SHAPE)), key
