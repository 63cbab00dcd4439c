"""FFTW_MEASURE reproducibility on the benchmark's shapes.

Each candidate is timed as the reference's timer does (kernel/timer.c:142-181:
the iteration count doubled until the minimum over 8 runs reaches the
minimum time) and the greedy search keeps a change only when it is more than
2% faster, so planning the same problem twice, with the wisdom forgotten in
between, must reach the same decisions."""
import pytest

import fftw3_amd as fa

pytestmark = pytest.mark.gpu


def decisions(p):
    # kernel, geometry and mapping of every pass; the batch layout follows
    return [" ".join(s.split(" lines=")[0].split()) for s in p.pass_info()]


def twice(make):
    out = []
    for _ in range(2):
        fa.lib().fftw_forget_wisdom()
        p = make()
        out.append(decisions(p))
        p.destroy()
    fa.lib().fftw_forget_wisdom()
    return out


def test_measure_reproducible_c2c_4096x1024_fp64(eng):
    n, H = 4096, 1024
    a = fa.DeviceBuffer(n * H * 16)
    b = fa.DeviceBuffer(n * H * 16)
    d1, d2 = twice(lambda: fa.plan_many_dft("double", [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1,
                                            fa.FFTW_MEASURE))
    print(d1, d2)
    assert d1 == d2


def test_measure_reproducible_r2c_2048sq_x64_fp32(eng):
    n, H = 2048, 64
    nh = n // 2 + 1
    a = fa.DeviceBuffer(n * n * H * 4)
    b = fa.DeviceBuffer(n * nh * H * 8)
    d1, d2 = twice(lambda: fa.plan_many_dft_r2c("float", [n, n], H, a.ptr, None, 1, n * n, b.ptr, None, 1,
                                                n * nh, fa.FFTW_MEASURE))
    print(d1, d2)
    assert d1 == d2


def test_measure_reproducible_c2c3d_512_fp64(eng):
    n = 512
    a = fa.DeviceBuffer(n ** 3 * 16)
    d1, d2 = twice(lambda: fa.plan_many_dft("double", [n] * 3, 1, a.ptr, None, 1, n ** 3, a.ptr, None, 1,
                                            n ** 3, -1, fa.FFTW_MEASURE))
    print(d1, d2)
    assert d1 == d2
