"""The C-ABI library loads and exports every entry point include/lumina_hip.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from lumina_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "lumina_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lh_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_match_binding():
    assert header_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_binding_loads():
    lib = _lib.load()
    assert lib.lh_last_error() == b""


def test_struct_layout_matches_c(tmp_path):
    """sizeof/offsetof of the header's structs, as gcc lays them out, equal
    the ctypes mirrors used by the Python binding."""
    import subprocess
    checks = {
        "lh_c64": _lib.c64, "lh_device_info": _lib.DeviceInfo,
        "lh_pairsum_spec": _lib.PairsumSpec, "lh_incident": _lib.Incident,
        "lh_solve_info": _lib.SolveInfo, "lh_batch_problem": _lib.BatchProblem,
        "lh_kernel_times": _lib.KernelTimes,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lumina_hip.h"',
             "int main(void) {"]
    for cname, ct in checks.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for cname, ct in checks.items():
        assert int(out[cname]) == ctypes.sizeof(ct), cname
        for f, _ in ct._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(ct, f).offset, (cname, f)
