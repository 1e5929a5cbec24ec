"""The C-ABI library loads on a CPU-only host and exports every entry point
declared in include/gisapr_hip.h, with a matching ctypes binding."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gisapr_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gis_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def library():
    from gisapr_amd import _lib
    if not _lib._LIB_PATH.exists():
        from gisapr_amd.build import build
        build()
    return _lib.lib()


def test_header_declares_the_path():
    names = declared_functions()
    for n in ("gis_splat", "gis_volume_dft", "gis_build_templates", "gis_build_template_coefficients",
              "gis_prepare_particles",
              "gis_correlate", "gis_score_sum", "gis_last_error"):
        assert n in names


def test_every_declared_symbol_is_exported_and_bound(library):
    from gisapr_amd import _lib
    for name in declared_functions():
        assert name in _lib.SIGNATURES, name
        assert hasattr(library, name), name
        assert getattr(ctypes.CDLL(str(_lib._LIB_PATH)), name) is not None


def test_version_and_argument_errors_without_gpu(library):
    assert library.gis_version() == 1
    # argument validation happens before any device call
    rc = library.gis_splat(None, None, 0, 0, 6, 1.0, 1.0, 1.0, None, None)
    assert rc == -1 and b"box" in library.gis_last_error()
    rc = library.gis_correlate(None, None, None, None, None, None, None, None, 1, 5, 64, 34, 1, 33, 33, 0, 0, None,
                               None, None, None, None, None, None, None, 0, None)
    assert rc == -2 and b"crop" in library.gis_last_error()
    rc = library.gis_correlate_split(None, None, None, None, None, None, None, None, 1, 5, 64, 34, 1, 33, 33, 0, 0,
                                     None, None, None, None, None, None, None, None, 0, None)
    assert rc == -2 and b"crop" in library.gis_last_error()
    rc = library.gis_correlate_split(None, None, None, None, None, None, None, None, 1, 5, 64, 16, 1, 32, 33, 0, 0,
                                     None, None, None, None, None, None, None, None, 0, None)
    assert rc == -1 and b"row strides" in library.gis_last_error()
    rc = library.gis_correlate_split_f16(None, None, None, None, None, None, None, None, None, None, 1, 5, 64, 34,
                                         1, 0, 0, 0, None, None, None, None, None, None, None, None, 0, None)
    assert rc == -2 and b"crop" in library.gis_last_error()
    rc = library.gis_correlate_split_f16(None, None, None, None, None, None, None, None, None, None, 1, 5, 64, 16,
                                         1, 0, 0, 2, None, None, None, None, None, None, None, None, 0, None)
    assert rc == -1
    assert library.gis_correlate_split_f16_workspace(256, 16, 10, 100) > library.gis_correlate_split_workspace(
        256, 16, 10)
    rc = library.gis_pack_fragments_f16(None, None, 1, 64, 0.0, None, None)
    assert rc == -1 and b"scale" in library.gis_last_error()
    rc = library.gis_move_chain(None, -1, None, None, 1, None, None)
    assert rc == -1 and b"n_atoms" in library.gis_last_error()
    assert library.gis_move_chain(None, 0, None, None, 4, None, None) == 0  # nothing to move
    rc = library.gis_volume_dft(None, 32, None, None, 0, None)
    assert rc == -1
    rc = library.gis_build_template_coefficients(None, 32, 32, None, 1, None, None, None, 0, 16, None, None, None,
                                                 0, None)
    assert rc == -1 and b"null" in library.gis_last_error()
    assert library.gis_volume_dft_bytes(256) == 16 * 256 * 256 * 65 * 8
    # N data planes x 2N rows x 33 groups of 8 kx (257 padded to 264)
    assert library.gis_volume_dft_workspace(256) == 256 * 512 * 264 * 8
