"""Wisdom and planner-patience semantics.

CPU part: the text format's contract, as in FFTW (kernel/planner.c:786-906):
entries round-trip through export/import, wisdom made for another kernel
catalog is rejected, malformed strings are rejected atomically (nothing is
imported), forget_wisdom clears everything.

GPU part: FFTW_MEASURE / FFTW_PATIENT autotune on the device and record
wisdom; FFTW_WISDOM_ONLY returns a plan only when wisdom exists
(api/apiplan.c:102-107); every tuned plan stays correct.
"""
import ctypes

import numpy as np
import pytest

import fftw3_amd as fa
import oracle


def export():
    L = fa.lib()
    p = L.fftw_export_wisdom_to_string()
    s = ctypes.cast(p, ctypes.c_char_p).value.decode()
    L._libc.free(p)
    return s


def catalog_tag(s):
    return s.split()[1]


def test_roundtrip_and_atomic_rejection():
    L = fa.lib()
    L.fftw_forget_wisdom()
    s0 = export()
    tag = catalog_tag(s0)
    assert s0.startswith("(fftw3-amd-gfx950-wisdom #x")
    good = "(fftw3-amd-gfx950-wisdom %s\n  (#x00000000deadbeef 1 3 0 2 1)\n  (#x0000000000000abc 2 0)\n)\n" % tag
    assert L.fftw_import_wisdom_from_string(good.encode()) == 1
    s1 = export()
    assert "(#x00000000deadbeef 1 3 0 2 1)" in s1 and "(#x0000000000000abc 2 0)" in s1
    # a partially valid string imports nothing
    bad = "(fftw3-amd-gfx950-wisdom %s\n  (#x0000000000000001 1 1 5)\n  (#x02 1 2 0\n)" % tag
    assert L.fftw_import_wisdom_from_string(bad.encode()) == 0
    assert "#x0000000000000001" not in export()
    # another catalog's wisdom is rejected
    other = good.replace(tag, "#x0123456789abcdef")
    L.fftw_forget_wisdom()
    assert L.fftw_import_wisdom_from_string(other.encode()) == 0
    assert export().count("(#x") == 0
    # re-import of an export restores it
    assert L.fftw_import_wisdom_from_string(s1.encode()) == 1
    assert export() == s1
    L.fftw_forget_wisdom()
    assert export() == s0


def test_file_roundtrip(tmp_path):
    L = fa.lib()
    L.fftw_forget_wisdom()
    tag = catalog_tag(export())
    s = "(fftw3-amd-gfx950-wisdom %s\n  (#x0000000000000042 1 2 1 0)\n)\n" % tag
    assert L.fftw_import_wisdom_from_string(s.encode()) == 1
    fn = str(tmp_path / "w.txt")
    L.fftw_export_wisdom_to_filename.argtypes = [ctypes.c_char_p]
    L.fftw_import_wisdom_from_filename.argtypes = [ctypes.c_char_p]
    assert L.fftw_export_wisdom_to_filename(fn.encode()) == 1
    L.fftw_forget_wisdom()
    assert L.fftw_import_wisdom_from_filename(fn.encode()) == 1
    assert "(#x0000000000000042 1 2 1 0)" in export()
    assert L.fftw_import_wisdom_from_filename(b"/nonexistent/wisdom") == 0
    L.fftw_forget_wisdom()


def test_large_entries_export_exactly():
    """Many entries with the most decisions and the largest accepted values
    export completely; decision values beyond the accepted bound (and
    beyond int range) are rejected at import, atomically."""
    L = fa.lib()
    L.fftw_forget_wisdom()
    tag = catalog_tag(export())
    ent = ["  (#x%016x 3 32 %s)" % (0x1000 + i, " ".join(["63"] * 32)) for i in range(200)]
    s = "(fftw3-amd-gfx950-wisdom %s\n%s\n)\n" % (tag, "\n".join(ent))
    assert L.fftw_import_wisdom_from_string(s.encode()) == 1
    out = export()
    assert out.count("(#x") == 200
    for e in ent:
        assert e.strip() in out
    assert out.rstrip().endswith(")")
    for big in ("64", "2147483647", "99999999999"):
        bad = "(fftw3-amd-gfx950-wisdom %s\n  (#x0000000000000077 1 2 0 %s)\n)\n" % (tag, big)
        assert L.fftw_import_wisdom_from_string(bad.encode()) == 0
    assert "#x0000000000000077" not in export()
    L.fftw_forget_wisdom()


def _run(prec, n, howmany, flags, seed=5):
    from gpu_util import c2c_many
    return c2c_many(prec, [n], howmany, -1, seed=seed, flags=flags)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,n,howmany", [("double", 4096, 64), ("float", 1024, 256),
                                            ("double", 512, 64), ("float", 65536, 4)])
def test_measure_records_wisdom_and_stays_correct(eng, prec, n, howmany):
    from gpu_util import TOL
    L = fa.lib()
    L.fftw_forget_wisdom()
    y, ref, _ = _run(prec, n, howmany, fa.FFTW_MEASURE)
    assert oracle.rel_l2(y, ref) <= TOL[prec]
    assert export().count("(#x") == 1
    y, ref, _ = _run(prec, n, howmany, fa.FFTW_PATIENT)
    assert oracle.rel_l2(y, ref) <= TOL[prec]


@pytest.mark.gpu
def test_wisdom_only(eng):
    from gpu_util import TOL
    L = fa.lib()
    L.fftw_forget_wisdom()
    n, H = 2048, 32
    buf = fa.DeviceBuffer(n * H * 16)
    out = fa.DeviceBuffer(n * H * 16)
    nn = (ctypes.c_int * 1)(n)
    # no wisdom yet: WISDOM_ONLY gives NULL
    assert not L.fftw_plan_many_dft(1, nn, H, buf.ptr, None, 1, n, out.ptr, None, 1, n, -1,
                                    fa.FFTW_WISDOM_ONLY)
    _run("double", n, H, fa.FFTW_MEASURE)
    s = export()
    L.fftw_forget_wisdom()
    assert not L.fftw_plan_many_dft(1, nn, H, buf.ptr, None, 1, n, out.ptr, None, 1, n, -1,
                                    fa.FFTW_WISDOM_ONLY)
    assert L.fftw_import_wisdom_from_string(s.encode()) == 1
    y, ref, _ = _run("double", n, H, fa.FFTW_WISDOM_ONLY)
    assert oracle.rel_l2(y, ref) <= TOL["double"]
    L.fftw_forget_wisdom()


@pytest.mark.gpu
def test_timelimit_zero_still_plans(eng):
    from gpu_util import TOL
    L = fa.lib()
    L.fftw_forget_wisdom()
    L.fftw_set_timelimit(0.0)
    try:
        y, ref, _ = _run("float", 4096, 16, fa.FFTW_EXHAUSTIVE)
        assert oracle.rel_l2(y, ref) <= TOL["float"]
    finally:
        L.fftw_set_timelimit(-1.0)
        L.fftw_forget_wisdom()
