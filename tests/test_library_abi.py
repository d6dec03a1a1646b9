"""The gfx950 library loads on a CPU-only host and exports every entry point
declared in include/quasar_mhd.h (no compute is run here)."""
import ctypes as C
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "quasar_mhd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qmhd_\w+)\s*\(", text)))


def test_header_declares_the_expected_surface():
    names = declared_functions()
    for required in ("qmhd_reconstruct", "qmhd_hlld_flux", "qmhd_ct_emf", "qmhd_rk_stage",
                     "qmhd_cfl_max_rate", "qmhd_solver_create", "qmhd_solver_step",
                     "qmhd_solver_cfl_limit", "qmhd_solver_run"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from quasar_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(declared_functions())
    assert lib.qmhd_version().decode().startswith("quasar-mhd")


def test_python_mirror_validates_like_the_reference():
    from quasar_amd import mhd
    import pytest
    with pytest.raises(ValueError):
        mhd.Grid2D(0, 4, 1.0, 1.0)
    with pytest.raises(ValueError):
        mhd.reconstruction_halo("weno9")
    assert mhd.reconstruction_order("mp7") == 7 and mhd.reconstruction_halo("mp5") == 3
    assert mhd.registered_reconstructions() == ["mp5", "mp7", "muscl_minmod"]
    g = mhd.Grid2D(8, 4, 1.0, 0.5, 0.0, 0.0, 4)
    assert g.storage_size() == 16 * 12 and g.index(0, 0) == 4 + 16 * 4
    assert g.x_at_cell_center(3) == 3.5 / 8
    _ = C
