"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/*.h declares; host-only helpers behave."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(em_[a-z0-9_]+)\s*\(", txt, re.M)
    return names


def test_library_exports_every_declared_symbol():
    import emergence_amd
    from emergence_amd import _lib

    lib = emergence_amd.load()
    declared = _declared()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTS)


def test_case_init_layout():
    import emergence_amd

    u = emergence_amd.case_init(0, (16, 12, 1), (0, 0, 0), (1, 1, 1), 6)
    assert u.shape == (28, 24, 1, 5)
    inner = u[6:22, 6:18, 0]
    assert np.all(inner[..., 0] > 0)
    assert np.all(u[:6, :, 0] == 0)  # ghosts left empty
    # NE quadrant state (1.5, 0, 0, 1.5) -> E = 1.5 / 0.4
    np.testing.assert_allclose(inner[-1, -1], [1.5, 0, 0, 0, 1.5 / 0.4])


def test_null_handle_is_an_error():
    import emergence_amd

    lib = emergence_amd.load()
    assert lib.em_solver_step(None, C.c_double(1.0)) == -1
    assert b"null" in lib.em_last_error()
