"""C-ABI checks that need no GPU: the library loads, exports every symbol include/pb3d.h declares,
and rejects bad arguments on the host before any launch."""
import ctypes
import os
import re

import pytest

from polybin3d_amd import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pb3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pb3d_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_the_binding_binds():
    assert declared_symbols() == sorted(C.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.pb3d_version() >= 10000


def test_struct_layouts_match_header():
    # sizes of the C structs as laid out by the host compiler (natural alignment)
    assert ctypes.sizeof(C.Grid) == 4 * 4 + 9 * 8
    assert ctypes.sizeof(C.ROp) == 16 + 32 + 32 + 8
    assert ctypes.sizeof(C.KOp) == 24 + 24 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 16 + 16
    assert ctypes.sizeof(C.SubMap) == ctypes.sizeof(C.KOp) + 8 + 16


def test_host_side_argument_validation():
    lib = C.lib()
    p = ctypes.c_void_p()
    # sizes neither the power-of-two passes nor the mixed-radix passes take: odd n2, an axis of 1,
    # n0 / n1 above 1024, n2 above 2048 or below 4
    for n in ([4, 4, 5], [1, 4, 4], [4, 1, 4], [4, 4, 2], [1025, 4, 4], [4, 2000, 4], [4, 4, 2050]):
        g = C.make_grid(n, [1, 1, 1], [0, 0, 0], [0, 0, 1], 0)
        assert lib.pb3d_plan_create(ctypes.byref(p), ctypes.byref(g)) == -2
        assert b"n0, n1 in [2, 1024]" in lib.pb3d_last_error()
    g = C.make_grid([8, 8, 8], [0, 1, 1], [0, 0, 0], [0, 0, 1], 0)
    assert lib.pb3d_plan_create(ctypes.byref(p), ctypes.byref(g)) == -1
    assert b"boxsize" in lib.pb3d_last_error()
    assert lib.pb3d_fft_r2c(None, None, None, None, 1, None, None, None) == -1
    assert lib.pb3d_fft_c2r(None, None, None, None, 1, None, None, None) == -1
    g = C.make_grid([8, 8, 8], [1, 1, 1], [0, 0, 0], [0, 0, 1], 0)
    assert lib.pb3d_bin_reduce(ctypes.byref(g), None, 0, None, 0, 0, 1, 1, None, None, 0, None, 1, None, None,
                               None) == -1
    dummy = ctypes.c_void_p(16)
    # nout above the limit / bad mode / bad ell are rejected before any launch
    assert lib.pb3d_bin_reduce(ctypes.byref(g), dummy, 0, None, 0, 0, 1, 17, None, None, 0, dummy, 1, dummy,
                               dummy, None) == -1
    assert lib.pb3d_bin_reduce(ctypes.byref(g), dummy, 0, None, 0, 0, 1, 1, None, None, 2, dummy, 1, dummy,
                               dummy, None) == -1
    bad = C.int_array([7])
    assert lib.pb3d_bin_reduce(ctypes.byref(g), dummy, 0, None, 0, 0, 1, 1, bad, None, 0, dummy, 1, dummy,
                               dummy, None) == -1
    assert lib.pb3d_triplet_sum(None, 1, 1, None, None, None, 1, 0, None, None, None) == -1
    assert lib.pb3d_kmap(ctypes.byref(g), dummy, dummy, 1, None, None) == -1
    k = C.kop(kind=9)
    assert lib.pb3d_kmap(ctypes.byref(g), dummy, dummy, 1, ctypes.byref(k), None) == -1
    assert lib.pb3d_triplet_work_size(0, 1, 0) == 0


def test_pointing_work_size():
    """The pointing chain's work maps keep only the z planes inside the band limits: nsub*batch
    sub-maps with pzi = the largest support[2] + 1 over the sub-maps, batch output maps with
    pzo = support[2] + 1 of the output op; n2/2+1 where one has no support."""
    lib = C.lib()
    g = C.make_grid([256, 256, 256], [1000.] * 3, [0, 0, 0], [0, 0, 1], 0)
    arr = (C.SubMap * 3)()
    for j, sz in enumerate((40, 63, 20)):
        arr[j].kop = C.kop(C.KOP_PLAIN, support=[sz, sz, sz])
    out63 = C.kop(C.KOP_PLAIN, support=[63, 63, 63])
    assert lib.pb3d_fft_pointing_work_size(ctypes.byref(g), 5, 3, arr, ctypes.byref(out63)) == \
        5 * 256 * 256 * (3 * 64 + 64)
    out70 = C.kop(C.KOP_PLAIN, support=[70, 70, 70])
    assert lib.pb3d_fft_pointing_work_size(ctypes.byref(g), 5, 1, arr, ctypes.byref(out70)) == \
        5 * 256 * 256 * (41 + 71)
    # no output band limit, or a support reaching the Nyquist plane: whole half-complex rows
    assert lib.pb3d_fft_pointing_work_size(ctypes.byref(g), 2, 2, arr, None) == 2 * 256 * 256 * (2 * 64 + 129)
    nyq = C.kop(C.KOP_PLAIN, support=[128, 128, 128])
    assert lib.pb3d_fft_pointing_work_size(ctypes.byref(g), 1, 1, arr, ctypes.byref(nyq)) == 256 * 256 * (41 + 129)
    assert lib.pb3d_fft_pointing_work_size(ctypes.byref(g), 1, 17, arr, None) == -1
    assert lib.pb3d_fft_pointing_work_size(None, 1, 1, arr, None) == -1


def test_python_errors_are_raised():
    with pytest.raises(ValueError):
        C.check(-1)


def test_struct_layouts_match_c_compiler(tmp_path):
    """The ctypes mirrors have the sizes and field offsets the C compiler gives include/pb3d.h."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no host C compiler")
    src = tmp_path / "sizes.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"pb3d.h\"\n"
                   "int main(){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(pb3d_grid), sizeof(pb3d_rop),"
                   " sizeof(pb3d_kop), sizeof(pb3d_submap), offsetof(pb3d_kop, support), offsetof(pb3d_kop, ft),"
                   " offsetof(pb3d_submap, coef));}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == [ctypes.sizeof(C.Grid), ctypes.sizeof(C.ROp), ctypes.sizeof(C.KOp), ctypes.sizeof(C.SubMap),
                   C.KOp.support.offset, C.KOp.ft.offset, C.SubMap.coef.offset]


def test_fisher_tiles_host_list():
    """pb3d_fisher_tiles (host only): tile count, largest segment first, row tiles of one column tile
    consecutive, odd or negative sizes rejected."""
    lib = C.lib()
    segs = []
    for k, rows in ((64, 130), (1000, 5), (0, 7), (200, 300)):
        s = C.FisherSeg()
        s.k, s.rows = k, rows
        segs.append(s)
    arr = (C.FisherSeg * 4)(*segs)
    nq = 300
    n = lib.pb3d_fisher_tiles(arr, 4, nq, None, 0)
    assert n == 3 * (2 + 1 + 3)
    tiles = (ctypes.c_int32 * (3 * n))()
    assert lib.pb3d_fisher_tiles(arr, 4, nq, tiles, n - 1) == -1
    assert lib.pb3d_fisher_tiles(arr, 4, nq, tiles, n) == n
    t = [tuple(tiles[3 * i:3 * i + 3]) for i in range(n)]
    assert [x[0] for x in t] == [1] * 3 + [3] * 9 + [0] * 6
    assert t[3:6] == [(3, 0, 0), (3, 128, 0), (3, 256, 0)]
    segs[0].k = 63
    arr = (C.FisherSeg * 4)(*segs)
    assert lib.pb3d_fisher_tiles(arr, 4, nq, None, 0) == -1
    assert b"odd" in lib.pb3d_last_error()

