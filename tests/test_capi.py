"""The C ABI library (built for gfx950) loads on a host without a GPU and
exports every entry point declared in include/*.h."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(tempo_hip_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for need in ("tempo_hip_init", "tempo_hip_run", "tempo_hip_run_host", "tempo_hip_finalize",
                 "tempo_hip_last_error", "tempo_hip_table_export", "tempo_hip_table_write"):
        assert need in names


def test_library_exports_every_declared_symbol():
    from tempo_amd import host
    lib = host.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_cfg_defaults_follow_reference():
    """tempo_hip_cfg_default = ty_tempo_cfgs defaults (src/module_mp_tempo_cfgs.F90:15-34);
    a host-only call, no device work."""
    from tempo_amd import host
    c = host.default_cfg()
    assert (c.aerosolaware_flag, c.hailaware_flag, c.semi_sedi_flag, c.turn_off_micro_flag) == (1, 1, 0, 0)
    assert (c.cloud_condensation_flag, c.refl10cm_flag, c.re_cloud_flag, c.max_hail_diameter_flag) == (1, 1, 1, 1)
    assert (c.rain_med_vol_diam_flag, c.graupel_med_vol_diam_flag, c.cloud_number_mixing_ratio_flag) == (0, 0, 0)
    assert c.nc_option == 0
    c = host.default_cfg(single_moment_nc_opt="10.e6")
    assert c.nc_option == 2 and c.nc_value == 10.e6


def test_library_is_built_for_gfx950():
    from tempo_amd import host
    data = open(host.LIB_PATH, "rb").read()
    assert b"gfx950" in data
