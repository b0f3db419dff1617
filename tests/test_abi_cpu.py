"""The C-ABI library without a GPU: it loads, exports every entry point that
include/lsgpu.h declares, agrees with the ctypes mirror on struct sizes, and
fills the reference's CLI defaults (src/main.cpp:39,89-91).  No compute call
is made: ls_create must refuse cleanly when no HIP device is present."""
import ctypes as C
from pathlib import Path

import pytest

from leachingsim_amd import lsgpu as L

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def lib():
    return L.load_library()


def test_every_declared_symbol_is_exported(lib):
    names = L.exported_symbols()
    assert len(names) >= 30
    assert "ls_net_num_weights" in names and "ls_step" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_binding(lib):
    sp, sr = C.c_size_t(), C.c_size_t()
    ver = lib.ls_abi_info(C.byref(sp), C.byref(sr))
    hdr = (ROOT / "include" / "lsgpu.h").read_text()
    assert f"#define LSGPU_ABI_VERSION {ver}" in hdr
    assert sp.value == C.sizeof(L.Params)
    assert sr.value == C.sizeof(L.StepReport)


def test_defaults_are_the_reference_cli_defaults(lib):
    p = L.Params()
    lib.ls_default_params(C.byref(p))
    ref = L.default_params()
    for name, _ in L.Params._fields_:
        assert getattr(p, name) == getattr(ref, name), name
    # src/main.cpp:39 option table and :89-91
    assert (p.D, p.dx, p.kreac, p.c_sat, p.top_pressure, p.max_cap_grad) == \
        (1.0, 1.0, 1.0, 0.95, 10000.0, 50000.0)
    assert (p.vg_alpha, p.vg_n, p.s_res, p.porosity, p.alpha_transfer) == \
        (1.0, 2.0, 0.05, 0.1, 1e-5)
    assert p.press_rtol == 1e-8 and p.press_maxit == 10000


def test_network_weight_count(lib):
    from leachingsim_amd import unet_spec as S
    lib.ls_net_num_weights.restype = C.c_size_t
    assert lib.ls_net_num_weights() == S.n_weights() == 366672


def test_create_without_device_fails_cleanly(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(L.LsError):
        L.GpuSimulation((8, 8, 8))


@pytest.mark.parametrize("field,value", [("mixed_schur", 4), ("mixed_schur", -1),
                                         ("adr_schur", 2), ("solver", 3),
                                         ("warm_history", 0), ("learned_race", -1.0)])
def test_create_rejects_out_of_range_settings(lib, field, value):
    """ls_create checks its settings before it touches the device, so the
    refusal (-3, no context returned) is the same with or without a GPU."""
    kw = {field: value}
    if field == "warm_history":
        kw["warm_start"] = L.WARM_PROJECTED
    p = L.default_params(**kw)
    h = C.c_void_p()
    assert lib.ls_create(C.byref(h), 8, 8, 8, 1, C.byref(p)) == -3
    assert not h.value
