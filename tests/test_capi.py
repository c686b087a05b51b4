"""CPU checks of the drop-in boundary: the library loads and exports every symbol the
headers in include/ declare (no compute calls: this host has no GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

import meshoptimizer_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in ("meshoptimizer.h", "meshopt_hip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        c_part = text.split("#if defined(__cplusplus) && !defined(MESHOPTIMIZER_NO_WRAPPERS)")[0]
        for m in re.finditer(r"\b(meshopt_\w+)\s*\(", c_part):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    if not os.path.exists(M.LIB_PATH):
        pytest.skip("libmeshopt_hip.so not built")
    lib = ctypes.CDLL(M.LIB_PATH)
    declared = _declared_symbols()
    assert declared == set(M.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_header_is_c89_clean(tmp_path):
    src = tmp_path / "ansi.c"
    src.write_text('#include "meshoptimizer.h"\n#include "meshopt_hip.h"\nint main(void) { return 0; }\n')
    res = subprocess.run(["gcc", "-std=c89", "-pedantic", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                          "-c", str(src), "-o", str(tmp_path / "ansi.o")], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_product_does_not_link_the_oracle():
    if not os.path.exists(M.LIB_PATH):
        pytest.skip("libmeshopt_hip.so not built")
    out = subprocess.run(["nm", "-D", M.LIB_PATH], capture_output=True, text=True).stdout
    assert "moracle" not in out
    ldd = subprocess.run(["ldd", M.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in ldd and "meshopt_ref" not in ldd


def test_index_adapters_compile(tmp_path):
    """The 16-bit index adapters (src/meshoptimizer.h:1514-1559) instantiate for unsigned short."""
    src = tmp_path / "adapters.cpp"
    src.write_text(
        '#include "meshoptimizer.h"\n'
        "size_t f(unsigned short* d, const unsigned short* i, float* v, unsigned char* l) {\n"
        "  size_t n = meshopt_simplify(d, i, 3, v, 3, 12, 3, 1e-2f);\n"
        "  n += meshopt_simplifyWithAttributes(d, i, 3, v, 3, 12, v, 12, v, 0, l, 3, 1e-2f);\n"
        "  n += meshopt_simplifyWithUpdate(d, 3, v, 3, 12, v, 12, v, 0, l, 3, 1e-2f);\n"
        "  n += meshopt_simplifySloppy(d, i, 3, v, 3, 12, l, 3, 1e-2f);\n"
        "  n += meshopt_simplifySloppy(d, i, 3, v, 3, 12, 3, 1e-2f);\n"
        "  return n;\n}\n")
    res = subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                          "-o", str(tmp_path / "adapters.o")], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_batch_rejects_bad_offsets_before_any_copy():
    """meshopt_hip_simplify_batch checks every host offset before it stages or launches anything:
    offsets that are not monotone (mesh 0 would read 3000 indices where 30 are staged), an index
    count that is not a multiple of 3, and a target above its mesh's index count are rejected with
    an error.  The checks precede every HIP call, so they run on a host without a GPU."""
    if not os.path.exists(M.LIB_PATH):
        pytest.skip("libmeshopt_hip.so not built")
    import numpy as np
    lib = M.library()
    idx = np.zeros(30, np.uint32)
    verts = np.zeros((8, 3), np.float32)
    dest = np.zeros(3000, np.uint32)
    counts = np.zeros(2, np.uint32)
    cases = [
        (np.array([0, 3000, 30], np.uint32), np.array([0, 4, 8], np.uint32), np.array([0, 0], np.uint32)),
        (np.array([0, 10, 30], np.uint32), np.array([0, 4, 8], np.uint32), np.array([0, 0], np.uint32)),
        (np.array([0, 15, 30], np.uint32), np.array([0, 4, 8], np.uint32), np.array([18, 0], np.uint32)),
        (np.array([0, 15, 30], np.uint32), np.array([0, 8, 4], np.uint32), np.array([0, 0], np.uint32)),
    ]
    for io, vo, tg in cases:
        rc = lib.meshopt_hip_simplify_batch(dest.ctypes.data, counts.ctypes.data, None, idx.ctypes.data, io.ctypes.data,
                                            verts.ctypes.data, vo.ctypes.data, None, 12, tg.ctypes.data, 2, 1e-2, 0)
        assert rc == -1
        assert "offsets" in lib.meshopt_hip_last_error().decode()


def test_batch_shape_validation():
    """meshopt_hip_batch_set_shape checks a launch shape on the host (no device call): widths of
    128/256/512/1024, the wavefronts per SIMD each width's registers allow, workgroups per CU not growing
    over the classes, the last class one workgroup per CU."""
    if not os.path.exists(M.LIB_PATH):
        pytest.skip("libmeshopt_hip.so not built")
    bad = [
        ([6, 5, 2, 1], [256, 256, 192, 1024], [0, 0, 0, 1], "128, 256, 512 or 1024"),
        ([7, 5, 2, 1], [256, 256, 256, 1024], [0, 0, 0, 1], "register budget"),
        ([13, 5, 2, 1], [128, 256, 256, 1024], [0, 0, 0, 1], "register budget"),
        ([4, 5, 2, 1], [256, 256, 256, 1024], [0, 0, 0, 1], "must not grow"),
        ([6, 5, 2, 2], [256, 256, 256, 256], [0, 0, 0, 1], "whole CU"),
        ([6, 5, 2, 1], [256, 256, 256, 1024], [0, 0, 0, 1][:3], "four values"),
    ]
    for per_cu, threads, cursor, why in bad:
        with pytest.raises(ValueError, match=why):
            M.batch_set_shape(per_cu, threads, cursor)
    M.batch_set_shape([5, 4, 2, 1], [256, 256, 512, 512], [0, 0, 1, 1])
    M.batch_set_shape([12, 9, 5, 1], [128, 128, 256, 1024], [0, 0, 0, 1])
    M.batch_set_shape()


ALLOC_SRC = os.path.join(ROOT, "tests", "capi", "alloc_test.cpp")


def build_alloc_test(out_dir):
    """Compiles the C++ allocator test against include/ and libmeshopt_hip.so (host code only)."""
    exe = os.path.join(str(out_dir), "alloc_test")
    res = subprocess.run(["g++", "-std=c++11", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), ALLOC_SRC, "-o", exe,
                          "-L", os.path.dirname(M.LIB_PATH), "-lmeshopt_hip", "-Wl,-rpath," + os.path.dirname(M.LIB_PATH)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


def test_cpp_overloads_build(tmp_path):
    """The header's C++ overloads and index adapters compile as C++11 and link against the library."""
    if not os.path.exists(M.LIB_PATH):
        pytest.skip("libmeshopt_hip.so not built")
    build_alloc_test(tmp_path)


@pytest.mark.gpu
def test_set_allocator_counts_adapter_blocks(tmp_path):
    """meshopt_setAllocator with a counting allocator (demo/tests.cpp:1712-1762): the 16-bit overloads
    take their adapter blocks through it and free every one before returning, 32-bit calls take
    none, the default allocator restored is the one used afterwards; 16-bit results equal 32-bit."""
    exe = build_alloc_test(tmp_path)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "alloc ok" in res.stdout, res.stdout + res.stderr
