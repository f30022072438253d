"""The C ABI library: loads on a CPU-only host, exports every entry point that
include/sad_hip.h declares, and rejects bad arguments with a status code and a
message before touching a device."""
import ctypes as C

import numpy as np
import pytest

from sad_amd import _lib, config


def test_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _lib.declared_symbols()
    assert len(declared) >= 28
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing
    assert set(_lib.SIGNATURES) <= set(declared)
    assert lib.sadh_version().decode().startswith("sad_hip")


@pytest.mark.parametrize("name,args", [
    ("sadh_render_voronoi", [None, None, None, None, 1.0, 1, 4, 4, 4, 4, None]),
    ("sadh_update_candidates", [None] * 5 + [1, 0, 1.0, 0, 64.0, 0, 4, None, None, 0, 0, 1, 4, 4, 4, 4, None]),
    ("sadh_radix_sort_pairs", [None, None, None, None, 16, 0xFFFFFFFF, None]),
    ("sadh_clear_buffer", [None, 4, None]),
])
def test_null_pointers_are_rejected(name, args):
    lib = _lib.load()
    rc = getattr(lib, name)(*args)
    assert rc == -1
    assert "null" in lib.sadh_last_error().decode() or "bad" in lib.sadh_last_error().decode()


def test_ssim_loss_needs_rendered():
    """The SSIM term reads `rendered`: a null one is refused before any launch."""
    lib = _lib.load()
    z = C.c_void_p(16)
    args = [z] * 15
    args[3] = None                                   # rendered
    rc = lib.sadh_compute_gradients_tiled(*args, z, 1.0, 1, z, 0, 0.5, 2, 2, 2, 2, None)
    assert rc == -1 and "rendered" in lib.sadh_last_error().decode()


def test_fit_plan_and_capacity_error():
    from sad_amd import trainer
    o = config.fit_options(n_sites=16384, target_bpp=1.0, iterations=2000, densify_freq=100, prune_freq=100,
                           max_sites=262144)
    p = trainer.fit_plan(o, 1024, 1024)
    assert p["capacity"] == 262144 and p["pair_count"] == 262144
    assert 0 < p["densify_percentile"] < p["prune_percentile"] < 0.95
    s = config.to_struct(o)
    sites = np.zeros((10, 10), np.float32)
    res = _lib.FitResult()
    rc = _lib.load().sadh_fit(C.byref(s), sites.ctypes.data, None, 1024, 1024, sites.ctypes.data, 10, None, 0,
                              C.byref(res))
    assert rc == -3 and "plan needs" in _lib.load().sadh_last_error().decode()


def test_module_map_imports_the_headers(tmp_path):
    """include/module.modulemap (the Swift/Clang-module view of the ABI) names headers that compile as a
    module: a C translation unit that imports it sees the status codes and the site layout."""
    import os
    import shutil
    import subprocess
    clang = shutil.which("clang") or "/opt/rocm/llvm/bin/clang"
    if not os.path.exists(clang):
        pytest.skip("no clang to read the module map")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src = tmp_path / "use.m"
    src.write_text("@import SadHip;\n_Static_assert(sizeof(SadSite) == 40, \"Site\");\n"
                   "int f(void) { return SADH_OK + SADH_ERR_CAPACITY; }\n")
    r = subprocess.run([clang, "-fmodules", f"-fmodules-cache-path={tmp_path}", "-fsyntax-only", "-I", inc, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of the option and result structs (sad_amd/_lib.py) have the header's size and field
    offsets, compiled from include/sad_hip.h by the host C compiler."""
    import os
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    structs = {"SadhFitOptions": _lib.FitOptions, "SadhFitResult": _lib.FitResult, "SadhBatchJob": _lib.BatchJob,
               "SadhBatchResult": _lib.BatchResult, "SadhServerOptions": _lib.ServerOptions}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sad_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run([cc, "-I", inc, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            parts = line.split()
            got[(parts[0], parts[1])] = int(parts[2])
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
