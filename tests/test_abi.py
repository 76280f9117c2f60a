"""The C ABI library loads and exports every entry point include/aldvc_hip.h declares.

No compute call is made (this runs on CPU-only hosts too); only the size queries and the
argument checks that return before any HIP call.
"""

from __future__ import annotations

import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "aldvc_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aldvc_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from pyaldvc_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} is not built (run __graft_entry__.build())")
    return _lib.lib()


def test_header_declares_the_binding_table():
    from pyaldvc_amd import _lib

    assert declared() == sorted(_lib.SIGNATURES)


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_header_compiles_as_c_and_cxx(lang, compiler):
    import shutil
    import subprocess

    if shutil.which(compiler) is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip(f"{compiler} or the ROCm headers are missing")
    res = subprocess.run([compiler, "-fsyntax-only", "-Wall", "-Werror", "-x", lang, "-D__HIP_PLATFORM_AMD__",
                          "-I/opt/rocm/include", HEADER], capture_output=True, text=True)  # fmt: skip
    assert res.returncode == 0, res.stderr[-2000:]


def test_library_exports_every_declared_symbol(lib):
    for name in declared():
        assert hasattr(lib, name), name


def test_abi_version_and_sizes(lib):
    from pyaldvc_amd import _lib

    assert lib.aldvc_abi_version() == _lib.ABI_VERSION
    # both variants: count + node list; plain adds 2048 scratch rows of S floats, fused adds
    # 32 float64 reference moments per node
    S = 33 ** 3
    b0 = lib.aldvc_icgn12_workspace_bytes(10000, 16, 16, 16, 1, 0)
    b1 = lib.aldvc_icgn12_workspace_bytes(10000, 16, 16, 16, 1, 1)
    assert b0 - b1 == 4 * 2048 * S - 256 * 10000
    assert b1 >= 8 * 10000 + 256 * 10000
    assert lib.aldvc_icgn3_workspace_bytes(10) >= 80


def _default_args(name):
    from pyaldvc_amd import _lib

    out = []
    for t in _lib.SIGNATURES[name][1]:
        if t is ctypes.c_void_p:
            out.append(None)
        elif t is ctypes.c_double:
            out.append(0.0)
        else:
            out.append(8)
    return out


def test_argument_errors_are_reported_without_a_gpu(lib):
    from pyaldvc_amd import _lib

    args = _default_args("aldvc_icgn12")
    args[16] = 7  # interp mode: rejected before any HIP call
    assert lib.aldvc_icgn12(*args) == -2
    assert b"interp mode" in lib.aldvc_last_error()
    args = _default_args("aldvc_zncc_map")  # null volumes
    assert lib.aldvc_zncc_map(*args) == -2 and b"null" in lib.aldvc_last_error()
    args = _default_args("aldvc_icgn3")
    args[19] = 1
    args[6] = -1  # negative half-width
    assert lib.aldvc_icgn3(*args) == -2 and b"half-widths" in lib.aldvc_last_error()
    with pytest.raises(_lib.HipLibraryError):
        _lib.check(-2, "probe")
    assert lib.aldvc_last_kernel_ms() == -1.0  # nothing timed on this thread


def test_library_carries_the_hash_of_the_sources_on_disk(lib):
    from pyaldvc_amd import build

    assert lib.aldvc_source_hash().decode() == build.source_hash()


def test_a_library_built_from_other_sources_is_refused(tmp_path, monkeypatch):
    import shutil

    from pyaldvc_amd import _lib, build

    # a copy of the sources with one byte changed: the loaded library no longer matches it
    csrc = tmp_path / "csrc"
    shutil.copytree(build.CSRC, csrc)
    info = csrc / "aldvc_build_info.hip"
    info.write_text(info.read_text() + "\n")
    assert build.source_hash(str(csrc)) != build.source_hash()
    monkeypatch.setattr(build, "CSRC", str(csrc))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.HipLibraryError, match="built from other sources"):
        _lib.lib()


def test_noalloc_entry_rejects_calls_beyond_its_scratch(lib):
    # 254,199 nodes fit the fixed 64 MiB scratch of the fused variants; one more does not. The size
    # check comes before any HIP call, so this runs without a GPU.
    args = _default_args("aldvc_icgn12_noalloc")
    args[1] = 254_200
    args[3] = args[4] = args[5] = 8
    args[6] = 1
    args[-2] = 1  # fused variant
    assert lib.aldvc_icgn12_noalloc(*args) == -2
    assert b"254,199 nodes" in lib.aldvc_last_error()
    assert lib.aldvc_icgn12_workspace_bytes(254_199, 8, 8, 8, 1, 1) <= 64 << 20
    assert lib.aldvc_icgn12_workspace_bytes(254_200, 8, 8, 8, 1, 1) > 64 << 20


def test_loop_descriptor_layout_matches_the_header(tmp_path):
    """The ctypes mirror of AldvcAdmmLoop has the C compiler's size and field offsets."""
    import shutil
    import subprocess

    from pyaldvc_amd._lib import AdmmLoopDesc

    if shutil.which("gcc") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("gcc or the ROCm headers are missing")
    names = [f[0] for f in AdmmLoopDesc._fields_]
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "aldvc_hip.h"', "int main(void) {",
             '  printf("%zu\\n", sizeof(AldvcAdmmLoop));']  # fmt: skip
    lines += [f'  printf("%zu\\n", offsetof(AldvcAdmmLoop, {n}));' for n in names]
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    res = subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{os.path.dirname(HEADER)}",
                          str(src), "-o", str(exe)], capture_output=True, text=True)  # fmt: skip
    assert res.returncode == 0, res.stderr[-2000:]
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got[0] == ctypes.sizeof(AdmmLoopDesc)
    assert got[1:] == [getattr(AdmmLoopDesc, n).offset for n in names]


def test_loop_entries_refuse_bad_descriptors_without_a_gpu(lib):
    from pyaldvc_amd._lib import AdmmLoopDesc

    D = AdmmLoopDesc()
    D.struct_bytes = ctypes.sizeof(AdmmLoopDesc) - 8  # a caller built against another layout
    assert lib.aldvc_admm_loop_enqueue(ctypes.byref(D), 0, None) == -2 and b"layout" in lib.aldvc_last_error()
    D.struct_bytes = ctypes.sizeof(AdmmLoopDesc)
    D.nx, D.ny, D.nz, D.n_nodes, D.max_steps = 4, 3, 2, 24, 3
    D.pcg_budget = D.pcg_limit = D.fill_budget = D.fill_limit = 10
    for name, _ in AdmmLoopDesc._fields_:  # dummy addresses: every check returns before any HIP call
        if name in ("U1", "U2", "v", "U1raw", "U1n", "U2n", "F2", "W", "F2n", "status", "zncc", "n_iter", "bad",
                    "history", "pcg_iterations", "fill_iterations", "stop", "steps", "capped"):  # fmt: skip
            setattr(D, name, 256)
    D.method = 2
    assert lib.aldvc_admm_loop_enqueue(ctypes.byref(D), 0, None) == -2 and b"method" in lib.aldvc_last_error()
    D.method = 1
    D.pcg_budget = 11
    assert lib.aldvc_admm_loop_enqueue(ctypes.byref(D), 0, None) == -2 and b"budgets" in lib.aldvc_last_error()
    D.pcg_budget = 10
    h = ctypes.c_void_p()
    assert lib.aldvc_admm_loop_graph(ctypes.byref(D), None, ctypes.byref(h)) == -2
    assert b"null stream" in lib.aldvc_last_error()
    assert lib.aldvc_graph_launch(None, None) == -2
    assert lib.aldvc_graph_destroy(None) == 0
