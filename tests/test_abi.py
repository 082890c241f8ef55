"""The C-ABI library: it loads without a GPU, exports every symbol that
include/hg.h declares, and its settings structs have the reference's std140
layout (glsl/bindings.glsl:39-68).  No compute calls (no GPU here)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

import hydrogen as hg
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "hg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hg.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(hg.EXPORTS) == names, "binding list out of date with include/hg.h"


def test_erosion_settings_std140_layout():
    S = hg.ErosionSettings
    assert ctypes.sizeof(S) == 96
    expect = {"particle_count": 0, "Kc": 4, "Kalpha": 8, "Kconv": 16, "Ks": 24, "Kd": 32, "Ke": 40,
              "energy_kept": 44, "Kspeed": 48, "G": 56, "d_t": 60, "density": 64, "init_volume": 68,
              "friction": 72, "inertia": 76, "min_volume": 80, "min_velocity": 84, "ttl": 88}
    for name, off in expect.items():
        assert getattr(S, name).offset == off, name
    # the oracle's independent copy of the struct agrees
    for name, off in expect.items():
        assert getattr(O.ErosionSettings, name).offset == off, name
    assert ctypes.sizeof(O.ErosionSettings) == 96


def test_ref_particle_std430_layout():
    """Particle (glsl/bindings.glsl:101-111) in std430: vec2 members 8-aligned."""
    P = hg.RefParticle
    assert ctypes.sizeof(P) == 48 and ctypes.sizeof(O.RefParticle) == 48
    offs = {"sc": 0, "iters": 4, "position": 8, "velocity": 16, "volume": 24, "sediment": 32, "to_kill": 40}
    for name, off in offs.items():
        assert getattr(P, name).offset == off, name
        assert getattr(O.RefParticle, name).offset == off, name


def test_rain_settings_std140_layout():
    R = hg.RainSettings
    assert ctypes.sizeof(R) == 32
    assert (R.amount.offset, R.mountain_thresh.offset, R.mountain_multip.offset, R.period.offset,
            R.drops.offset) == (0, 4, 8, 12, 16)


def test_map_settings_std140_layout():
    offs = dict(max_height=0, max_dirt=4, hmap_dims=8, height_mult=16, water_lvl=20, seed=24, persistance=28,
                lacunarity=32, scale=36, redistribution=40, octaves=44, fake_erosion=48, mask_round=52,
                mask_exp=56, mask_power=60, mask_slope=64, uplift=68, uplift_scale=72, domain_warp=76,
                domain_warp_scale=80, terrace=84, terrace_scale=88)
    for S in (hg.MapSettings, O.MapSettings):
        assert ctypes.sizeof(S) == 96
        for name, off in offs.items():
            assert getattr(S, name).offset == off, name


def test_create_without_device_fails_cleanly():
    # In a container without a GPU hg_create must fail with a message, not crash.
    import torch
    if torch.cuda.is_available():
        return
    try:
        hg.Terrain(64)
    except hg.HGError as e:
        assert e.code in (hg.HG_ERR_HIP, hg.HG_ERR_INVALID)
    else:
        raise AssertionError("hg_create succeeded without a device")


def test_missing_library_raises_instead_of_falling_back(monkeypatch, tmp_path):
    """Without the built HIP library the binding raises: there is no CPU path."""
    from hydrogen import _abi
    monkeypatch.setattr(_abi, "_lib", None)
    monkeypatch.setattr(_abi, "LIB_PATH", str(tmp_path / "libhydrogen.so"))
    try:
        _abi.load()
    except RuntimeError as e:
        assert "missing" in str(e)
    else:
        raise AssertionError("load() succeeded without the library")


@pytest.mark.parametrize("lang,std,cc", [("c", "c11", "gcc"), ("c++", "c++17", "g++")])
def test_header_is_plain_c_and_cxx(lang, std, cc, tmp_path):
    """include/hg.h compiles on its own as C11 and as C++17, warnings as errors:
    the boundary is a C ABI with no C++ or HIP types."""
    import shutil
    import subprocess
    if shutil.which(cc) is None:
        pytest.skip(f"{cc} not installed")
    hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "hg.h")
    r = subprocess.run([cc, f"-std={std}", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-x", lang, hdr],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
