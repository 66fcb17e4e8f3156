"""The C ABI library loads and exports every declared symbol (no GPU calls)."""
import ctypes
import os

import numpy as np

from gpuocean_amd import Common, _lib, host


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert "gpo_cdklm16_substep" in declared
    for name in declared:
        assert hasattr(L, name), name
    assert L.gpo_version().decode().startswith("gpuocean_amd")


def test_library_is_built_for_gfx950_only():
    with open(_lib.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


def test_bad_arguments_are_reported_without_a_gpu():
    L = _lib.lib()
    p = _lib.CDKLM16Params()
    p.nx, p.ny, p.rk_order = 0, 10, 2
    rc = L.gpo_cdklm16_substep(ctypes.byref(p), 1.0, 0, *([None] * 9), 256, None, 256, None, None, 0x01010101, 0,
                               None)
    assert rc == -1
    assert b"positive" in L.gpo_last_error()
    p.nx = 10
    rc = L.gpo_cdklm16_rk2_steps(ctypes.byref(p), 3, 1.0, *([None] * 6), 256, None, 256, None, 0x03010101, 0, None)
    assert rc == -1
    assert b"wall" in L.gpo_last_error()


def test_bc_packing_matches_reference_layout():
    bc = Common.BoundaryConditions(north=2, east=1, south=2, west=1)
    assert bc.packed() == (2 << 24) | (2 << 16) | (1 << 8) | 1
    s = "north: Periodic, east: Wall, south: Periodic, west: Wall, spongeCells: {'north': 0, 'south': 0, 'east': 0, 'west': 0}"
    assert Common.BoundaryConditions.fromstring(s).packed() == bc.packed()


def test_periodic_mismatch_asserts():
    import pytest
    with pytest.raises(AssertionError):
        Common.BoundaryConditions(north=2, south=1)


def test_kernel_constant_roundtrip():
    for x in (9.81, 200.0, 0.1, 1e-5, 1.8, 0.8, 1025.0):
        assert host.kernel_constant(x) == np.float32(x)


def test_subsample_texture_is_exact_for_linear_fields():
    y, x = np.meshgrid(np.arange(74) + 0.5, np.arange(54) + 0.5, indexing="ij")
    data = 0.01 + 1e-6 * 200 * y
    tex = host.subsample_texture(data, 10)
    ys = (np.arange(7) + 0.5) / 7 * 74
    assert np.allclose(tex, (0.01 + 1e-6 * 200 * ys)[None, :], rtol=0, atol=1e-12)


def test_oracle_is_not_imported_by_the_product():
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpuocean_amd")
    for dirpath, _, files in os.walk(root):
        for fn in files:
            if fn.endswith(".py"):
                with open(os.path.join(dirpath, fn)) as f:
                    text = f.read()
                assert "from oracle" not in text and "import oracle" not in text, fn


def test_noalloc_adjoint_sweep_rejects_oversized_grids_without_a_gpu():
    """The size check of the no-allocation sweep comes before any device call."""
    L = _lib.lib()
    p = _lib.CDKLM16Params()
    p.nx, p.ny, p.rk_order = 2000, 2000, 2
    lam = (ctypes.c_void_p * 3)(None, None, None)
    rc = L.gpo_cdklm16_adjoint_steps_f32_noalloc(ctypes.byref(p), 1, 1.0, lam, lam, lam, 4 * 2048, None, 4 * 2048,
                                                  None, 0x01010101, None, None, None)
    assert rc == -1
    assert b"scratch" in L.gpo_last_error()
