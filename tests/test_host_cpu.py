"""The C++ host side without a GPU: the lsim command line, the .raw reader
(src/MPIRawLoader.h:48-82), the VTK image writer (src/Simulation.cpp:1306-1566,
src/main.cpp:510-526) decoded back value by value, and the checkpoint formats
(own, and the reference's one-rank layout, src/Simulation.cpp:2235-2307)."""
import ctypes as C
import re
import struct
import subprocess
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import pytest

PKG = Path(__file__).resolve().parents[1] / "leachingsim_amd"
LSIM = PKG / "lsim"
HOSTLIB = PKG / "liblshost.so"

VTK_ALL = 0x7FD
NAMES = {"MaterialType", "Concentration", "Pressure", "SulphideFrac", "Flux", "Saturation",
         "CapillaryPressure", "RelativePermeability", "EffectivePermeability", "AcidConcentration"}


@pytest.fixture(scope="module")
def lib():
    if not HOSTLIB.exists():
        pytest.fail("liblshost.so not built: run __graft_entry__.build()")
    h = C.CDLL(str(HOSTLIB))
    h.lshost_last_error.restype = C.c_char_p
    h.lshost_read_raw.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_void_p]
    h.lshost_write_vtk.restype = C.c_longlong
    h.lshost_write_vtk.argtypes = [C.c_char_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_uint, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
    h.lshost_write_pvd.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_size_t),
                                   C.c_size_t, C.c_char_p]
    h.lshost_write_checkpoint.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                          C.POINTER(C.c_void_p)]
    h.lshost_load_checkpoint.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return h


def ptrs(arrays):
    return (C.c_void_p * len(arrays))(*[a.ctypes.data if a is not None else None for a in arrays])


def test_lsim_help_lists_the_reference_options():
    out = subprocess.run([str(LSIM), "--help"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    for opt in ("raw-file", "xext", "yext", "zext", "out_dir", "header_size", "voxel_size", "D",
                "kext", "kreac", "Dpore_fac", "dt", "tmax", "nout", "seed", "csat", "evap_flux",
                "theta", "instant_precip", "porosity", "top_pressure", "max_cap_grad",
                "implicit_sat", "restart_step", "chk_freq", "debug", "use_out_time", "out_time",
                "time_scalar", "enable_microporosity", "rock_D_fac", "alpha_transfer"):
        assert re.search(rf"--{re.escape(opt)}\s", out.stdout), opt


def test_lsim_missing_required_option_and_no_device(tmp_path):
    out = subprocess.run([str(LSIM), "--xext", "4"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "required" in out.stderr
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    raw = tmp_path / "g.raw"
    np.full(4 * 4 * 4, 2, np.uint8).tofile(raw)
    out = subprocess.run([str(LSIM), "--raw-file", str(raw), "--xext", "4", "--yext", "4",
                          "--zext", "4", "--out_dir", str(tmp_path / "o")],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "HIP device" in out.stderr


@pytest.mark.parametrize("lb", [1, 2])
def test_raw_reader_skips_header_and_rejects_short_files(lib, tmp_path, lb):
    g = (3, 4, 5)
    dt = np.uint8 if lb == 1 else np.uint16
    data = (np.arange(np.prod(g)) % 5 + 1).astype(dt)
    f = tmp_path / "x.raw"
    f.write_bytes(b"HEADER" * 3 + data.tobytes())
    out = np.empty(np.prod(g), dt)
    assert lib.lshost_read_raw(str(f).encode(), *g, lb, 18, out.ctypes.data) == 0
    assert np.array_equal(out, data)
    assert lib.lshost_read_raw(str(f).encode(), *g, lb, 40, out.ctypes.data) == -1
    assert b"short" in lib.lshost_last_error()


def parse_vti(path):
    raw = Path(path).read_bytes()
    head, _, rest = raw.partition(b"<AppendedData encoding=\"raw\">")
    data = rest[rest.index(b"_") + 1:]
    root = ET.fromstring(head.decode() + "</VTKFile>" if b"</VTKFile>" not in head else head)
    return root, data


def read_block(data, offset, dtype):
    (nb,) = struct.unpack_from("<Q", data, offset)
    return np.frombuffer(data, dtype=dtype, count=nb // np.dtype(dtype).itemsize,
                         offset=offset + 8)


@pytest.mark.parametrize("lb", [1, 2])
def test_vtk_writer_names_types_and_values(lib, tmp_path, lb):
    ni, nj, nk = 5, 4, 3
    n = ni * nj * nk
    rng = np.random.default_rng(lb)
    img = (rng.integers(1, 6, n)).astype(np.uint8 if lb == 1 else np.uint16)
    frac = rng.random(n).astype(np.float32)
    f64 = [rng.standard_normal(n) for _ in range(12)]
    f64[2][7] = np.nan                                     # pressure: scrubbed to 0
    f64[2][11] = np.inf
    root = str(tmp_path / "output")
    nan = lib.lshost_write_vtk(root.encode(), 12, ni, nj, nk, lb, VTK_ALL, img.ctypes.data,
                               frac.ctypes.data, ptrs(f64))
    assert nan == 2
    # master file
    m = ET.parse(tmp_path / "output_000012.pvti").getroot()
    pim = m.find("PImageData")
    assert pim.get("WholeExtent") == f"0 {ni - 1} 0 {nj - 1} 0 {nk - 1}"
    arrays = {a.get("Name"): a for a in pim.find("PPointData")}
    assert set(arrays) == NAMES
    assert arrays["MaterialType"].get("type") == ("UInt8" if lb == 1 else "UInt16")
    assert arrays["SulphideFrac"].get("type") == "Float32"
    assert arrays["Flux"].get("NumberOfComponents") == "3"
    assert pim.find("Piece").get("Source") == "output_0_000012.vti"
    # the piece: decode every appended block; VTK point order is i fastest
    root_el, data = parse_vti(tmp_path / "output_0_000012.vti")
    das = {d.get("Name"): d for d in root_el.iter("DataArray")}
    assert set(das) == NAMES

    def vtk_order(a, comps=1):
        return np.asarray(a).reshape(ni, nj, nk, comps).transpose(2, 1, 0, 3).reshape(-1)

    types = {"UInt8": np.uint8, "UInt16": np.uint16, "Float32": np.float32, "Float64": np.float64}
    got = {k: read_block(data, int(d.get("offset")), types[d.get("type")]) for k, d in das.items()}
    p = f64[2].copy()
    p[~np.isfinite(p)] = 0.0
    assert np.array_equal(got["MaterialType"], vtk_order(img))
    assert np.array_equal(got["SulphideFrac"], vtk_order(frac))
    assert np.array_equal(got["Pressure"], vtk_order(p))
    assert np.array_equal(got["Concentration"], vtk_order(f64[3]))
    assert np.array_equal(got["AcidConcentration"], vtk_order(f64[4]))
    assert np.array_equal(got["Saturation"], vtk_order(f64[1]))
    assert np.array_equal(got["Flux"], vtk_order(np.stack(f64[5:8], -1), 3))
    assert np.array_equal(got["RelativePermeability"], vtk_order(f64[8]))
    assert np.array_equal(got["CapillaryPressure"], vtk_order(f64[9]))
    assert np.array_equal(got["EffectivePermeability"], vtk_order(f64[10]))


def test_pvd_time_mapping(lib, tmp_path):
    t = (C.c_double * 3)(0.0, 1.5, 3.0)
    s = (C.c_size_t * 3)(0, 1, 2)
    assert lib.lshost_write_pvd(str(tmp_path / "time_mapping.pvd").encode(), t, s, 3, b"output") == 0
    root = ET.parse(tmp_path / "time_mapping.pvd").getroot()
    ds = root.find("Collection").findall("DataSet")
    assert [d.get("file") for d in ds] == ["output_000000.pvti", "output_000001.pvti",
                                           "output_000002.pvti"]
    assert [float(d.get("timestep")) for d in ds] == [0.0, 1.5, 3.0]


def state(rng, n, lb):
    img = rng.integers(1, 6, n).astype(np.uint8 if lb == 1 else np.uint16)
    return img, rng.random(n).astype(np.float32), [rng.standard_normal(n) for _ in range(5)]


@pytest.mark.parametrize("lb", [1, 2])
def test_checkpoint_round_trip_and_reference_layout(lib, tmp_path, lb):
    g = (4, 3, 5)
    n = int(np.prod(g))
    rng = np.random.default_rng(7)
    img, frac, f64 = state(rng, n, lb)
    d = str(tmp_path).encode()
    assert lib.lshost_write_checkpoint(d, 40, 12.5, 0.5, *g, lb, img.ctypes.data,
                                       frac.ctypes.data, ptrs(f64)) == 0
    img2, frac2 = np.empty_like(img), np.empty_like(frac)
    out = [np.empty(n) for _ in range(5)]
    t, dt = C.c_double(), C.c_double()
    assert lib.lshost_load_checkpoint(d, 40, *g, lb, img2.ctypes.data, frac2.ctypes.data,
                                      ptrs(out), C.byref(t), C.byref(dt)) == 1
    assert (t.value, dt.value) == (12.5, 0.5)
    assert np.array_equal(img2, img) and np.array_equal(frac2, frac)
    assert all(np.array_equal(a, b) for a, b in zip(out, f64))
    # the reference's one-rank file: the 7 arrays back to back
    img, frac, f64 = state(rng, n, lb)
    (tmp_path / "checkpoint_rank0_000041.bin").write_bytes(
        img.tobytes() + frac.tobytes() + b"".join(a.tobytes() for a in f64))
    assert lib.lshost_load_checkpoint(d, 41, *g, lb, img2.ctypes.data, frac2.ctypes.data,
                                      ptrs(out), C.byref(t), C.byref(dt)) == 0
    assert np.array_equal(img2, img) and np.array_equal(frac2, frac)
    assert all(np.array_equal(a, b) for a, b in zip(out, f64))
    # wrong grid: size check
    assert lib.lshost_load_checkpoint(d, 41, 4, 3, 4, lb, img2.ctypes.data, frac2.ctypes.data,
                                      ptrs(out), C.byref(t), C.byref(dt)) == -1
    assert lib.lshost_load_checkpoint(d, 42, *g, lb, img2.ctypes.data, frac2.ctypes.data,
                                      ptrs(out), C.byref(t), C.byref(dt)) == -1


# ---------------------------------------------------------------------------
# The reference's per-solve log lines (src/Simulation.cpp:457,838,1278,1772,
# 2018,2208) as lsim prints them, read back with the regexes of the
# reference's benchmark parser (plot_benchmarks.py:22-24, restated): it sums
# every "Sub-steps: N" of an explicit-saturation log (the [Saturation] and
# [Leaching] lines) and every "[Implicit Saturation] Converged in N".
# ---------------------------------------------------------------------------
REGEX_EXPLICIT_STEPS = re.compile(r"Sub-steps:\s+(\d+)")
REGEX_IMPLICIT_STEPS = re.compile(r"\[Implicit Saturation\] Converged in\s+(\d+)")


def step_log(lib, rep, dt, params):
    from leachingsim_amd import lsgpu as L
    lib.lshost_step_log.argtypes = [C.POINTER(L.StepReport), C.c_double, C.POINTER(L.Params),
                                    C.c_char_p, C.c_size_t]
    buf = C.create_string_buffer(4096)
    n = lib.lshost_step_log(C.byref(rep), dt, C.byref(params), buf, len(buf))
    assert 0 < n < len(buf)
    return buf.value.decode()


def parse_like_reference(text, explicit):
    total = 0
    for line in text.splitlines():
        m = (REGEX_EXPLICIT_STEPS if explicit else REGEX_IMPLICIT_STEPS).search(line)
        if m:
            total += int(m.group(1))
    return total


def test_step_log_lines_match_reference_formats(lib):
    from leachingsim_amd import lsgpu as L
    P = L.default_params()
    r = L.StepReport(press_iters=17, press_relres=3e-9, p_min=0.0, p_max=2.5, sat_substeps=7,
                     sat_max_divq=0.445, adr_iters=5, adr_iters_conc=5, adr_iters_acid=0,
                     adr_relres=1e-9, conc_min=0.0, conc_max=1.0, acid_min=0.0, acid_max=0.0,
                     leach_substeps=3, leach_max_rate=0.25)
    text = step_log(lib, r, 0.5, P)
    lines = text.splitlines()
    assert lines[0] == "    [Pressure] Converged in 17 iterations. (Reason Code: 2)"
    assert lines[1] == "    [Pressure] Min Value: 0 Pa | Max Value: 2.5 Pa"
    assert lines[2] == "    [Saturation] Max div(q): 0.445 | Sub-steps: 7 (0.0714286s)"
    assert lines[3] == "    [Saturation] Sub-stepping required: 7 steps of 0.0714286s."
    assert lines[4] == "    [Concentration] Converged in 5 iterations. (Reason Code: 2)"
    assert lines[6] == "    [Acid Transport] Converged in 0 iterations. (Reason Code: 2)"
    assert lines[8] == "    [Leaching] Max rate: 0.25 | Sub-steps: 3 (0.166667s)"
    assert parse_like_reference(text, explicit=True) == 7 + 3
    # implicit saturation: the parser counts the solver's iterations instead
    Pi = L.default_params(implicit_sat=1)
    r.sat_iters, r.sat_valid, r.sat_raw_min, r.sat_raw_max = 9, 1, -0.01, 0.93
    text = step_log(lib, r, 0.5, Pi)
    assert "    [Implicit Saturation] Converged in 9 iterations. (Reason Code: 2)" in text
    assert parse_like_reference(text, explicit=False) == 9
    assert "[Saturation] Max div(q)" not in text
    # a pressure solve that missed the tolerance ends the step's lines
    r.press_relres = 1e-3
    text = step_log(lib, r, 0.5, P)
    assert text == "    [Pressure] DIVERGED/FAILED in 17 iterations! (Error Code: -3)\n"
