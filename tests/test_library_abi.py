"""The C-ABI library loads and exports every symbol include/lentils_hip.h
declares; the by-value struct layouts match the reference's ctypes mirrors
(load_backend.py:27-57).  No GPU call is made."""
import ctypes as C
import os
import re

import numpy as np

from lentils_amd import backend as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, 'include', 'lentils_hip.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    names = re.findall(r'^\s*(?:void|lt_status|long|const char \*)\s*\*?\s*(\w+)\s*\(', text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(B.LIBPATH)
    declared = _declared_functions()
    assert len(declared) >= 30
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert set(B.exported_symbols()) <= set(declared)


def test_struct_layouts_match_reference_abi():
    assert C.sizeof(B.image_space_ctype) == 4 * 4 + 6 * 8 + 2 * 8
    assert C.sizeof(B.fourier_space_ctype) == 8 * 3 + 4 * 8 + 8  # 5 ints padded to 24 bytes
    assert C.sizeof(B.visibility_space_ctype) == 16 + 16
    assert C.sizeof(B.global_lens_model_ctype) == 16
    assert B.generic_mass_model_ctype.itemsize == 320
    assert [f for f, _ in B.image_space_ctype._fields_][:4] == ['nx', 'ny', 'num_channels', 'num_stokes']


def test_binding_table_loads():
    L = B.lib()
    assert L.lt_version().decode().startswith('lentils_hip')
