"""The C ABI library loads, exports every function include/geoswe_abi.h declares, and the
ctypes structures match the C layout (checked against gcc's offsetof on the header)."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from geoswe_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
HDR = ROOT / "include" / "geoswe_abi.h"


def declared_functions():
    txt = re.sub(r"/\*.*?\*/", "", HDR.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gs_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    h = C.CDLL(str(L.LIB_PATH))
    missing = [n for n in names if not hasattr(h, n)]
    assert not missing, missing
    assert set(names) == set(L._SIGS), set(names) ^ set(L._SIGS)


def test_version_and_error_string():
    assert L.lib().gs_version() == 1
    rc = L.lib().gs_dense_step(None, None, L.GsForcing(), None, None)
    assert rc == 1 and b"NULL" in L.lib().gs_last_error()


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layout_matches_header(tmp_path):
    structs = {"GsClock": L.GsClock, "GsPhys": L.GsPhys, "GsForcing": L.GsForcing, "GsDense": L.GsDense,
               "GsFlat": L.GsFlat, "GsAdjFlat": L.GsAdjFlat, "GsFlatMeshHost": L.GsFlatMeshHost}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HDR}"', "int main(void){"]
    for sname, cls in structs.items():
        lines.append(f'printf("{sname} %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{sname}.{fname} %zu\\n", offsetof({sname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = dict(l.split() for l in out if l.strip())
    for sname, cls in structs.items():
        assert int(want[sname]) == C.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert int(want[f"{sname}.{fname}"]) == getattr(cls, fname).offset, (sname, fname)


def test_new_entry_points_reject_bad_arguments():
    """Argument checks of the ensemble and adjoint entry points run before any device call."""
    lib = L.lib()
    out = C.c_void_p()
    assert lib.gs_ens_create(None, None, None, None, 1, C.byref(out)) == 1
    assert b"NULL" in lib.gs_last_error()
    assert lib.gs_ens_create(None, None, None, None, 0, C.byref(out)) == 3     # GS_E_RANGE
    assert lib.gs_ens_step(None, None) == 1
    assert lib.gs_ens_set_launch(None, 0, 1) == 1
    assert lib.gs_ens_destroy(None) == 0
    assert lib.gs_adj_flat_step(None, None, C.c_double(1.0), None) == 1
    a = L.GsAdjFlat()
    a.n_stored = 10
    assert lib.gs_adj_flat_step(C.byref(a), C.byref(L.GsPhys()), C.c_double(1.0), None) == 1
    assert b"NULL" in lib.gs_last_error()
    assert lib.gs_adj_add(None, 0, None, None, 0, None) == 0                  # nothing to add
    assert lib.gs_adj_add(None, 0, None, None, -1, None) == 1
    assert lib.gs_flat_replay_pre(None, None, L.GsForcing(), None, 1, None, None, None, None) == 1
    assert lib.gs_flat_run_frozen(None, None, L.GsForcing(), None, 1, None, None, 0, None, 1, None, None) == 1
    assert lib.gs_adj_flat_nsum(None, None) == 1
