"""The C-ABI library loads and exports every entry point declared in include/svimx.h."""

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "svimx.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s*\*?\s*(svimx_[a-z0-9_]+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "svimx_fractions" in syms and "svimx_nr_exec" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol():
    from svierosion_amd import _abi
    lib = ctypes.CDLL(str(_abi.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    bound = set(_abi.exported_symbols())
    assert not [s for s in declared_symbols() if s not in bound]


def test_error_mapping_without_gpu():
    from svierosion_amd import _abi
    assert _abi.lib.svimx_version() >= 1
    assert _abi.lib.svimx_workspace_bytes(90, 90, 90) > 4096
    # a null workspace is rejected on the host before any device call
    rc = _abi.lib.svimx_fractions(None, None, None, 0, 10, 10, 10, 0.1, None, None, 5.5, None, 0)
    assert rc == _abi.EBADARG
    assert "workspace" in _abi.lib.svimx_last_error().decode()


def test_nr_plan_refuses_asymmetric_weights_without_gpu():
    """The N_r paths pair the windows of +dx and -dx (and rely on symmetric z
    taps and a real spectrum): a weight table that is not mirror-symmetric is
    refused on the host, before any device call; a ball's table passes that
    check (the plan then needs the GPU)."""
    import numpy as np
    from svierosion_amd import _abi
    from oracle.flow import ball_cell_weights
    w = ball_cell_weights(0.3, 0.05)
    m = (w.shape[0] - 1) // 2
    assert np.array_equal(w.astype(np.float32), w[::-1].astype(np.float32))
    assert np.array_equal(w.astype(np.float32), w[:, ::-1].astype(np.float32))
    assert np.array_equal(w.astype(np.float32), w[:, :, ::-1].astype(np.float32))
    bad = np.ascontiguousarray(w.copy())
    bad[m + 1, m, m] += 0.25  # breaks the x mirror only
    h = ctypes.c_void_p()
    rc = _abi.lib.svimx_nr_plan_create(bad.ctypes.data_as(ctypes.c_void_p), m, 0.05, 0.3, 40, 40, 40,
                                       ctypes.byref(h))
    assert rc == _abi.EBADARG and not h.value
    assert "mirror-symmetric" in _abi.lib.svimx_last_error().decode()


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """ctypes mirrors of the header's structs have the C compiler's offsets and sizes."""
    import shutil
    import subprocess
    from svierosion_amd import _abi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no host C compiler")
    structs = {"svimx_law_t": _abi.Law, "svimx_job_desc_t": _abi.JobDesc, "svimx_step_record_t": _abi.StepRec,
               "svimx_job_status_t": _abi.JobStatus, "svimx_reinit_gate_t": _abi.ReinitGate}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "svimx.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
