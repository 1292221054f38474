"""The C ABI: libqss.so loads without a GPU, exports every function declared in
include/qss.h, and the ctypes mirror of qss_params has the C layout."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qss.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char)\s*\*?\s*(qss_\w+)\s*\(", txt,
                                 re.M)))


def test_header_declares_the_sweep_surface():
    names = declared_functions()
    for must in ("qss_create", "qss_destroy", "qss_sweep", "qss_layer_step", "qss_guess",
                 "qss_iterate", "qss_push_deposit", "qss_transform2d", "qss_beam_layer"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from tertia_amd import _lib
    L = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(declared_functions()) <= set(_lib.SIGNATURES), \
        set(declared_functions()) - set(_lib.SIGNATURES)


def test_create_without_gpu_fails_cleanly():
    from tertia_amd import _lib
    from tertia_amd.params import small
    L = _lib.load()
    c = small().to_c()
    h = ctypes.c_void_p()
    rc = L.qss_create(ctypes.byref(c), 0, 0, ctypes.byref(h))
    if rc == 0:          # a GPU is present: nothing to check here
        L.qss_destroy(h)
        pytest.skip("GPU present")
    assert rc in (-1, -2) and not h.value


@pytest.mark.parametrize("kw", [dict(precision=7), dict(precision=1, ny=24),
                                dict(precision=2, fix_subcycle=1)])
def test_create_rejects_bad_precision_before_touching_a_gpu(kw):
    """Parameter validation runs first, so these fail with QSS_E_ARG on any host."""
    from tertia_amd import _lib
    from tertia_amd.params import small
    L = _lib.load()
    c = small(**kw).to_c()
    h = ctypes.c_void_p()
    assert L.qss_create(ctypes.byref(c), 0, 0, ctypes.byref(h)) == -1
    assert b"precision" in L.qss_last_error()


def test_variant_tables():
    from tertia_amd import _lib
    L = _lib.load()
    for prec in range(4):
        assert L.qss_gemm_variants(prec) > 0
    assert L.qss_gemm_variants(9) == -1


def test_params_layout_matches_c():
    from tertia_amd.params import QssBunch, QssParams
    fields = [f for f, _ in QssParams._fields_]
    bfields = [f for f, _ in QssBunch._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "qss.h"', "int main(void){",
           'printf("%zu %zu\\n", sizeof(qss_params), sizeof(qss_bunch));']
    for f in fields:
        src.append('printf("%%zu\\n", offsetof(qss_params, %s));' % f)
    for f in bfields:
        src.append('printf("%%zu\\n", offsetof(qss_bunch, %s));' % f)
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(v) for v in out]
    assert vals[0] == ctypes.sizeof(QssParams) and vals[1] == ctypes.sizeof(QssBunch)
    for i, f in enumerate(fields):
        assert vals[2 + i] == getattr(QssParams, f).offset, f
    for i, f in enumerate(bfields):
        assert vals[2 + len(fields) + i] == getattr(QssBunch, f).offset, f


def test_driver_layout_matches_c():
    from tertia_amd.params import QssDriver
    fields = [f for f, _ in QssDriver._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "qss.h"', "int main(void){",
           'printf("%zu\\n", sizeof(qss_driver));']
    for f in fields:
        src.append('printf("%%zu\\n", offsetof(qss_driver, %s));' % f)
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(v) for v in out]
    assert vals[0] == ctypes.sizeof(QssDriver)
    for i, f in enumerate(fields):
        assert vals[1 + i] == getattr(QssDriver, f).offset, f


def test_driver_entry_points_refuse_before_seeding_without_a_gpu():
    """The driver and hand-off entry points validate their arguments first."""
    from tertia_amd import _lib
    L = _lib.load()
    assert L.qss_driver_push(None) == -1
    assert L.qss_time_step(None, None) == -1
    assert L.qss_driver_seed(None, None, None) == -1
    n = ctypes.c_size_t()
    assert L.qss_layer_state_bytes(None, ctypes.byref(n)) == -1
    assert L.qss_pack_layer(None, None) == -1
    assert L.qss_unpack_layer(None, None, 0) == -1
    assert L.qss_layer_record_doubles(None, ctypes.byref(n)) == -1
    assert L.qss_pack_layer_record(None, 0, None) == -1
    assert L.qss_unpack_layer_record(None, None, 0) == -1


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run(["python", os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_noalloc_transform_checks_sizes_before_touching_a_device():
    """qss_transform2d_noalloc refuses planes larger than its fixed scratch
    (512 x 512) and unknown kinds with QSS_E_ARG before any HIP call."""
    from tertia_amd import _lib
    L = _lib.load()
    assert L.qss_transform2d_noalloc(600, 32, 0, 0, None, None, 1.0, None) == -1
    assert b"512" in L.qss_last_error()
    assert L.qss_transform2d_noalloc(32, 32, 2, 0, None, None, 1.0, None) == -1
    assert L.qss_transform2d_noalloc(32, 32, 0, 0, None, None, 1.0, None) == -1

