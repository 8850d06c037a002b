"""The C-ABI library loads and exports every entry point its header declares
(no GPU needed: nothing is computed)."""
import os
import re

import pytest

from puffin_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "puffin_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(puff_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("puff_create", "puff_rk4par", "puff_diffraction_step", "puff_transform",
              "puff_upload_particles", "puff_download_field"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gpu.LIB_PATH):
        pytest.fail("libpuffin_gpu.so is not built (run `make -C puffin_amd`)")
    lib = gpu.lib()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    # the Python front end binds exactly what the header declares
    assert set(gpu.SIGNATURES) == set(declared_symbols())


def test_params_struct_layout_matches_header():
    import ctypes
    from puffin_amd.params import PuffParams
    text = open(HEADER).read()
    body = text.split("typedef struct puff_params {", 1)[1].split("} puff_params;", 1)[0]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ, rest = decl.split(None, 1)
        names += [n.strip() for n in rest.split(",")]
    assert names == [f[0] for f in PuffParams._fields_]
    assert ctypes.sizeof(PuffParams) == 16 * 4 + 19 * 8


H5_HEADER = os.path.join(ROOT, "include", "puffin_h5.h")


def test_h5_library_exports_every_declared_symbol():
    from puffin_amd import h5io
    text = re.sub(r"/\*.*?\*/", "", open(H5_HEADER).read(), flags=re.S)
    syms = sorted(set(re.findall(r"\b(puff_h5_[a-z0-9_]+)\s*\(", text)))
    assert len(syms) == 10
    lib = h5io.lib()
    for s in syms:
        assert hasattr(lib, s), s


def test_h5_info_struct_layout_matches_header():
    import ctypes
    from puffin_amd.h5io import H5Info
    text = open(H5_HEADER).read()
    body = text.split("typedef struct puff_h5_info {", 1)[1].split("} puff_h5_info;", 1)[0]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ, rest = decl.split(None, 1)
        names += [n.strip() for n in rest.split(",")]
    assert names == [f[0] for f in H5Info._fields_]
    assert ctypes.sizeof(H5Info) == 7 * 8 + 9 * 4 + 8 * 4 + 16 * 8 + 4
