"""The C ABI library: it loads, exports every entry point include/odevec.h
declares, and reports errors through return codes (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from odevec_amd import _lib

HDR = os.path.join(os.path.dirname(__file__), "..", "include", "odevec.h")


def declared_functions():
    text = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|void|const char\*)\s+(odevec_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIBPATH)
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names


def test_registry_and_network_info():
    nets = _lib.networks()
    for base in ("ROBER/NONE/DENSE", "OREGO/NONE/DENSE", "PRIMORDIAL/CMK/CSC", "PRIMORDIAL/FF/CSC"):
        assert base in nets and base + "+REPRO" in nets
    info = _lib.network_info("PRIMORDIAL/CMK/CSC")
    assert info["neq"] == 16 and info["nnz"] == 89 and info["nrate"] == 38 and info["tgas_index"] == 14
    assert _lib.network_info("ROBER/NONE/DENSE")["nparam"] == 3
    with pytest.raises(_lib.OdevecError) as e:
        _lib.network_info("NOPE/NONE/DENSE")
    assert e.value.code == _lib.ENOTFOUND
    assert _lib.lib().odevec_abi_version() == 2


def test_create_argument_checks():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.odevec_create(b"NOPE", 4, None, C.byref(h)) == _lib.ENOTFOUND
    assert L.odevec_create(b"ROBER/NONE/DENSE", 0, None, C.byref(h)) == _lib.EINVAL
    bad = _lib.Opts(7, 0, 0, 0, 0, 0, 0, 0)
    assert L.odevec_create(b"ROBER/NONE/DENSE", 4, C.byref(bad), C.byref(h)) == _lib.EINVAL
    grp = _lib.Opts(0, 1, 999, 0, 0, 0, 0, 0)
    assert L.odevec_create(b"ROBER/NONE/DENSE", 4, C.byref(grp), C.byref(h)) == _lib.EINVAL
    sched = _lib.Opts(0, _lib.PERSISTENT, 0, 0, 0, 0, 0, 0, 2)
    assert L.odevec_create(b"ROBER/NONE/DENSE", 4, C.byref(sched), C.byref(h)) == _lib.EINVAL
    assert L.odevec_set_tolerances(None, 1e-6, 1e-10) == _lib.EINVAL
    assert L.odevec_set_fp64_fallback(None, 128) == _lib.EINVAL
    assert L.odevec_set_invariant_test(None, 3.0, 32) == _lib.EINVAL
    assert L.odevec_set_mixed_convergence_test(None, 1) == _lib.EINVAL
    assert L.odevec_destroy(None) == _lib.OK


def test_create_without_device_fails_cleanly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = _lib.lib().odevec_create(b"ROBER/NONE/DENSE", 4, None, C.byref(h))
    assert rc == _lib.EDEVICE and not h.value
    assert "device" in _lib.last_error()


def test_fortran_module_links_and_fails_cleanly_without_gpu():
    """The Fortran module (bind(C) interfaces to this ABI) and a driver in the
    reference's calling pattern compile and link; without a GPU,
    InitOdevec reports the library's error and stops, as the reference's
    failures do."""
    from conftest import gpu_available
    from odevec_amd import build as B
    if not B.flang():
        pytest.skip("no Fortran compiler")
    if not os.path.exists(B.FORTRAN_DRIVER):
        B.build_fortran(verbose=False)
    if gpu_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([B.FORTRAN_DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "InitOdevec failed" in r.stdout and "no HIP device" in r.stdout


def test_bundles_registered_and_batch_argument_checks():
    assert "MIX3" in _lib.bundles() and "MIX3+REPRO" in _lib.bundles()
    ent = _lib.bundle_entries("MIX3")
    assert ent == [(n, p) for n in ("ROBER/NONE/DENSE", "OREGO/NONE/DENSE", "PRIMORDIAL/CMK/CSC") for p in (0, 1)]
    L = _lib.lib()
    h = C.c_void_p()
    assert L.odevec_batch_create(b"NOPE", 0, 0, 0, C.byref(h)) == _lib.ENOTFOUND
    assert L.odevec_batch_create(b"MIX3", 100, 0, 0, C.byref(h)) == _lib.EINVAL
    assert L.odevec_batch_solve_async(None, None, 0, None) == _lib.EINVAL
    assert L.odevec_batch_destroy(None) == _lib.OK


def test_cooperative_mode_registry_and_checks():
    info = _lib.network_info("SYNTH64/FF/CSC")
    assert info["neq"] == 64 and info["nparam"] == 1 and info["nrate"] == 300
    L = _lib.lib()
    h = C.c_void_p()
    lane = _lib.Opts(0, _lib.PERCELL, 0, 0, 0, 0, 0, 0)
    coop = _lib.Opts(0, _lib.COOP, 0, 0, 0, 0, 0, 0)
    assert L.odevec_create(b"SYNTH64/FF/CSC", 4, C.byref(lane), C.byref(h)) == _lib.EINVAL
    assert "cooperative" in _lib.last_error()
    assert L.odevec_create(b"ROBER/NONE/DENSE", 4, C.byref(coop), C.byref(h)) == _lib.EINVAL


def test_header_is_c99_and_links(tmp_path):
    """include/odevec.h is plain C: a C99 host compiles against it and links
    libodevec.so (every declared entry point referenced)."""
    import re
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "odevec.h")).read()
    names = sorted(set(re.findall(r"\b(odevec_[a-z_0-9]+)\s*\(", hdr)))
    src = tmp_path / "host.c"
    src.write_text('#include "odevec.h"\n#include <stdio.h>\nint main(void) {\n'
                   "  void* p[] = {%s};\n  odevec_opts_t o; odevec_default_opts(&o);\n"
                   '  printf("%%d %%d\\n", odevec_abi_version(), (int)(sizeof p / sizeof p[0]));\n  return 0;\n}\n'
                   % ", ".join("(void*)&" + n for n in names))
    exe = tmp_path / "host"
    libdir = os.path.join(root, "odevec_amd")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-pedantic-errors", "-Wno-pedantic",
                    "-I", os.path.join(root, "include"), str(src), "-o", str(exe), "-L", libdir, "-lodevec",
                    "-Wl,-rpath," + libdir], check=True)
