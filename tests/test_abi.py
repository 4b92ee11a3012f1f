"""The C ABI library: it loads without a GPU and exports every entry point
include/wig.h declares."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "wig.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wig_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("wig_create", "wig_destroy", "wig_fluid_step", "wig_report", "wig_upload", "wig_download",
                 "wig_sync", "wig_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from worldinagrain_amd import gpu
    lib = gpu.lib()                       # loads on a host with no GPU
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (wig_\w+)", out))
    assert set(declared_functions()) <= exported


def test_library_holds_gfx950_code():
    from worldinagrain_amd import gpu
    data = open(gpu.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_errors_without_device_are_reported():
    from worldinagrain_amd import gpu
    lib = gpu.lib()
    if lib.wig_device_ok():
        pytest.skip("a device is visible")
    with pytest.raises(gpu.WigError):
        gpu.GpuWorld(16, 16)


def test_air_errors_are_reported():
    # the air grid's refusals: a side out of range, no device; each NULL
    # with the reason in wig_last_error
    from worldinagrain_amd import air, gpu
    L = air.lib()
    assert not L.wig_air_create(0, 16)
    assert b"sides" in L.wig_last_error()
    assert not L.wig_air_create(16, 20000)
    assert L.wig_air_project(None, 1) < 0
    if not L.wig_device_ok():
        assert not L.wig_air_create(16, 16)
        assert L.wig_last_error()
        with pytest.raises(gpu.WigError):
            air.GpuAir(16, 16)
