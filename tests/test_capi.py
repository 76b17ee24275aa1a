"""The C ABI library loads without a GPU and exports every entry point include/rmesh.h declares;
argument checking that needs no device memory reports errors through rm_last_error."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "rmesh.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rm_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from radiance_meshes_amd import _capi
    path = _capi.LIB_PATH
    if not os.path.exists(path):
        from radiance_meshes_amd import build
        build.build()
    return _capi.lib()


def test_every_declared_symbol_is_exported(lib):
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_the_header():
    from radiance_meshes_amd import _capi
    assert set(_declared()) == set(_capi.SIGNATURES)


def test_argument_errors_without_device(lib):
    from radiance_meshes_amd import _capi
    cam = _capi.RmCamera()
    cam.tile_width = cam.tile_height = 5
    cam.image_height = cam.image_width = 20
    cam.grid_height = cam.grid_width = 4
    rc = lib.rm_composite_fwd(cam, None, 0, None, None, None, None, 7, None, None, None, None, None)
    assert rc == 2 and b"tile" in lib.rm_last_error()
    rc = lib.rm_mlp_fwd(None, 10, None, None, None, None, None, None, 40, 0, None, None)
    assert rc == 1 and b"n_out" in lib.rm_last_error()
    assert lib.rm_sort_pairs_workspace_size(1000) > 1000 * 12
    assert lib.rm_scan_workspace_size(1 << 20) > 0
    assert lib.rm_bin_workspace_size(1000, 5000, 64) > 5000 * 16


def test_hashgrid_workspace_rule(lib):
    """Slot-ownership backward applies to the finest run of hashed levels with res >= 512, at most
    8 of them (n_feat 2 only)."""
    from radiance_meshes_amd.field import level_table
    res, offs = level_table(16, 16, 1.38, 19)          # levels 11..15 (res 553..2005) qualify
    n = 1_000_000
    ws = lib.rm_hashgrid_bwd_workspace_size(n, res.ctypes.data, offs.ctypes.data, 16, 2, 19)
    n_part = sum(1 for r in res if int(r) ** 3 > 2 ** 19 and r >= 512)
    assert n_part == 5
    assert ws >= n * 8 * n_part * 10                  # (slot, value pair) per update
    assert ws < n * 8 * (n_part + 1) * 10             # and no level below res 512
    res24, offs24 = level_table(24, 16, 1.38, 19)      # 13 levels of res >= 512: the finest 8 qualify
    ws24 = lib.rm_hashgrid_bwd_workspace_size(n, res24.ctypes.data, offs24.ctypes.data, 24, 2, 19)
    assert sum(1 for r in res24 if r >= 512) == 13
    assert n * 8 * 8 * 10 <= ws24 < n * 8 * 9 * 10
    assert lib.rm_hashgrid_bwd_workspace_size(n, res.ctypes.data, offs.ctypes.data, 16, 4, 19) == 0
    res8, offs8 = level_table(8, 16, 1.38, 19)         # finest resolution 106: no level qualifies
    assert lib.rm_hashgrid_bwd_workspace_size(n, res8.ctypes.data, offs8.ctypes.data, 8, 2, 19) == 0
