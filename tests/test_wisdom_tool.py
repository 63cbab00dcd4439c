"""fftw-wisdom pre-planning tool (the reference's tools/fftw-wisdom.c).

CPU part: the problem-string grammar of libbench2/problem.c:229-318 as the
tool accepts it (type, placement, direction, split prefix, geometry, r2r
kinds after each dim, *howmany), the reference's canonical set (tools/fftw-wisdom.c:134-147 and
:291-310: 8 types per 1-D size, in-place types only for multi-dimensional
sizes, 300 problems), ordering by size, and rejection of what it cannot
plan.  GPU part: the wisdom it writes makes FFTW_WISDOM_ONLY planning
succeed in a fresh process, for both precisions."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "fftw3_amd", "fftw-wisdom")

need_tool = pytest.mark.skipif(not os.path.exists(TOOL), reason="fftw-wisdom is not built")


def run(*args, **kw):
    return subprocess.run([TOOL] + list(args), capture_output=True, text=True, timeout=kw.pop("timeout", 60),
                          **kw)


@need_tool
def test_problem_grammar():
    r = run("-l", "cof1024", "rib256x256", "/cib64x64*4", "rof1000*3", "c-7", "c+7")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    # sorted by total points
    assert [l.split()[0] for l in lines] == ["c-7", "c+7", "cof1024", "rof1000*3", "/cib64x64*4",
                                            "rib256x256"]
    d = {l.split()[0]: l for l in lines}
    assert d["cof1024"] == "cof1024 complex out-of-place forward rank=1 n=1024 howmany=1"
    assert d["rib256x256"] == "rib256x256 real in-place backward rank=2 n=256x256 howmany=1"
    assert d["/cib64x64*4"] == "/cib64x64*4 complex in-place backward split rank=2 n=64x64 howmany=4"
    assert "forward" in d["c-7"] and "backward" in d["c+7"]


@need_tool
@need_tool
def test_r2r_grammar():
    # libbench2/problem.c:152-195: a kind letter after each dim, R2HC default
    r = run("-l", "ko1024e10", "ki64x32o11*3", "k8", "ki16hx8b", "k5o00")
    assert r.returncode == 0, r.stderr
    d = {l.split()[0]: l for l in r.stdout.strip().splitlines()}
    assert d["ko1024e10"] == "ko1024e10 r2r out-of-place rank=1 n=1024 kind=redft10 howmany=1"
    assert d["ki64x32o11*3"] == "ki64x32o11*3 r2r in-place rank=2 n=64x32 kind=r2hc,rodft11 howmany=3"
    assert d["k8"].endswith("kind=r2hc howmany=1")
    assert "kind=dht,hc2r" in d["ki16hx8b"]
    assert "kind=rodft00" in d["k5o00"]


@need_tool
@pytest.mark.parametrize("bad", ["cof12q", "cof", "c0", "cofx4", "/rof64", "k64e2", "k64o111", "/k64",
                                 "cof8*", "rof64e10"])
def test_rejects(bad):
    r = run("-l", bad)
    assert r.returncode != 0 and "cannot parse" in r.stderr


@need_tool
def test_canonical_set():
    r = run("-c", "-l")
    assert r.returncode == 0
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 300
    multi = [l for l in lines if "x" in l.split()[0]]
    assert multi and all(" in-place " in l for l in multi)
    assert sum(1 for l in lines if l.startswith("cof1048576 ")) == 1


@need_tool
def test_help_and_version():
    r = run("-h")
    assert r.returncode == 0 and "Size syntax" in r.stdout
    r = run("-V")
    assert r.returncode == 0 and "fftw-wisdom" in r.stdout


CHECK = r"""
import sys
sys.path.insert(0, %(root)r)
import fftw3_amd as fa
fa.init()
L = fa.lib()
assert L.fftw_import_wisdom_from_filename(%(path)r.encode()) == 1
prec = %(prec)r
b = fa.DeviceBuffer(1 << 22)
o = fa.DeviceBuffer(1 << 22)
fl = fa.FFTW_ESTIMATE | fa.FFTW_WISDOM_ONLY
p = fa.plan_many_dft(prec, [1024], 1, b.ptr, None, 1, 1024, o.ptr, None, 1, 1024, -1, fl)
q = fa.plan_many_dft_r2c(prec, [64, 64], 1, b.ptr, None, 1, 64 * 66, b.ptr, None, 1, 64 * 33, fl)
k = fa.plan_many_r2r(prec, [512], 4, b.ptr, None, 1, 512, o.ptr, None, 1, 512, [fa.FFTW_REDFT10], fl)
try:
    fa.plan_many_dft(prec, [2048], 1, b.ptr, None, 1, 2048, o.ptr, None, 1, 2048, -1, fl)
    raise SystemExit("no wisdom for 2048 expected")
except RuntimeError:
    pass
print("ok")
"""


@pytest.mark.gpu
@need_tool
@pytest.mark.parametrize("prec", ["double", "float"])
def test_wisdom_only_after_tool(tmp_path, prec):
    out = str(tmp_path / "w.txt")
    args = ["-m", "-n", "-o", out, "cof1024", "rif64x64", "ko512e10*4"]
    if prec == "float":
        args.insert(0, "-s")
    r = run(*args, timeout=120)
    assert r.returncode == 0, r.stderr
    txt = open(out).read()
    assert txt.startswith("(fftw3-amd-gfx950-wisdom #x") and txt.count("(#x") == 3
    r = subprocess.run([sys.executable, "-c", CHECK % {"root": ROOT, "path": out, "prec": prec}],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
