"""Pins the CPU oracle (oracle/liboracle.so) before anything is checked
against it.

The reference's own tests store no golden vectors (tests/check.pl and
libbench2 generate problems and self-check them), and the reference library
cannot be built here (its build needs generated codelets and a generated
config.h).  So the oracle is pinned by:
  * an independent library routine: numpy.fft (pocketfft);
  * an independent brute-force O(n^2) DFT in long double with exact index
    reduction (shares no code with the FFT recursion);
  * closed forms and the libbench2 invariants (libbench2/verify-lib.c:
    impulse :284-325, linearity :327-356, time/frequency shift :360-414),
    Parseval and round trip;
  * known-answer values of splitmix64 (the input generator).
Each check is chosen so that a plausible mistake (sign, index, missing
twiddle, wrong Hermitian half) fails it; test_planted_mistakes_are_caught
shows that for the sign and Hermitian cases.
"""
import numpy as np
import pytest

SIZES = list(range(1, 101)) + [128, 243, 256, 512, 1000, 1001, 2187, 4093, 4096, 8191]


def crand(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


@pytest.mark.parametrize("n", SIZES)
def test_c2c_vs_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = crand(rng, n)
    assert orc.rel_l2(orc.fftn(x, -1), np.fft.fft(x)) < 2e-15
    assert orc.rel_l2(orc.fftn(x, +1), np.fft.ifft(x) * n) < 2e-15


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 16, 17, 60, 97, 101, 211, 256, 1001, 4093])
def test_c2c_vs_bruteforce(orc, n):
    rng = np.random.default_rng(1000 + n)
    x = crand(rng, n)
    for sign in (-1, 1):
        assert orc.rel_l2(orc.fftn(x, sign), orc.dft_direct(x, sign)) < 5e-16


def test_double_path_matches_long_double(orc):
    rng = np.random.default_rng(7)
    for n in (64, 1000, 4093, 65536):
        x = crand(rng, n)
        assert orc.rel_l2(orc.fftn(x, -1, prec=orc.D), orc.fftn(x, -1)) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 8, 12, 97, 1024, 4093])
def test_closed_forms(orc, n):
    k = np.arange(n)
    # impulse at 0 -> all ones; constant -> n at k = 0
    e0 = np.zeros(n, complex)
    e0[0] = 1
    assert np.max(np.abs(orc.fftn(e0) - 1)) < 1e-15
    c = orc.fftn(np.ones(n, complex))
    ref = np.zeros(n, complex)
    ref[0] = n
    assert np.max(np.abs(c - ref)) < 1e-12 * n
    # impulse at j0 -> exp(-2 pi i j0 k / n) (forward sign -1)
    j0 = n // 3
    ej = np.zeros(n, complex)
    ej[j0] = 1
    want = np.exp(-2j * np.pi * ((j0 * k) % n) / n)
    # (tolerance covers numpy's own rounding of the angle 2*pi*m/n)
    assert np.max(np.abs(orc.fftn(ej, -1) - want)) < 4e-15
    assert np.max(np.abs(orc.fftn(ej, +1) - np.conj(want))) < 4e-15
    # single frequency exp(+2 pi i f j / n) -> n * delta(k - f) forward
    f = (2 * n) // 5 if n > 1 else 0
    x = np.exp(2j * np.pi * ((f * k) % n) / n)
    y = orc.fftn(x, -1)
    ref = np.zeros(n, complex)
    ref[f] = n
    assert np.max(np.abs(y - ref)) < 1e-12 * n


@pytest.mark.parametrize("n", [6, 64, 100, 1001])
def test_linearity_shift_parseval_roundtrip(orc, n):
    rng = np.random.default_rng(n + 5)
    x, y = crand(rng, n), crand(rng, n)
    a, b = complex(0.3, -1.2), complex(-0.7, 0.4)
    X, Y = orc.fftn(x), orc.fftn(y)
    assert orc.rel_l2(orc.fftn(a * x + b * y), a * X + b * Y) < 1e-15
    s = 7 % n
    k = np.arange(n)
    # time shift: x[j - s] -> W^{s k} X[k]
    xs = np.roll(x, s)
    assert orc.rel_l2(orc.fftn(xs), X * np.exp(-2j * np.pi * ((s * k) % n) / n)) < 1e-14
    # frequency shift: x[j] W^{-s j} -> X[k - s]
    assert orc.rel_l2(orc.fftn(x * np.exp(2j * np.pi * ((s * k) % n) / n)), np.roll(X, s)) < 1e-14
    assert abs(np.vdot(X, X).real - n * np.vdot(x, x).real) < 1e-12 * n * np.vdot(x, x).real
    assert orc.rel_l2(orc.fftn(X, +1) / n, x) < 1e-15


@pytest.mark.parametrize("shape", [(4, 6), (3, 5, 7), (8, 8, 8), (2, 3, 4, 5)])
def test_multid_vs_numpy(orc, shape):
    rng = np.random.default_rng(len(shape))
    x = crand(rng, *shape)
    assert orc.rel_l2(orc.fftn(x, -1), np.fft.fftn(x)) < 2e-15
    assert orc.rel_l2(orc.fftn(x, +1), np.fft.ifftn(x) * x.size) < 2e-15


def test_guru_strides_and_batches(orc):
    """FFTW layout semantics: element (j, k) of howmany k at j*stride + k*dist
    (doc/reference.texi:1062-1070); strides need not be contiguous."""
    rng = np.random.default_rng(11)
    n, H, stride, dist = 12, 5, 3, 40
    buf = crand(rng, 400)
    out = np.zeros(400, complex)
    orc.guru_dft([(n, stride, 2)], [(H, dist, 31)], buf, out, -1)
    for h in range(H):
        ref = np.fft.fft(buf[h * dist + stride * np.arange(n)])
        assert orc.rel_l2(out[h * 31 + 2 * np.arange(n)], ref) < 2e-15
    # in-place with identical strides
    b2 = buf.copy()
    orc.guru_dft([(n, stride, stride)], [(H, dist, dist)], b2, b2, +1)
    for h in range(H):
        ref = np.fft.ifft(buf[h * dist + stride * np.arange(n)]) * n
        assert orc.rel_l2(b2[h * dist + stride * np.arange(n)], ref) < 2e-15


@pytest.mark.parametrize("shape", [(1,), (2,), (7,), (16,), (100,), (1001,), (6, 9), (4, 5, 8), (3, 1, 6)])
def test_r2c_c2r_vs_numpy(orc, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.uniform(-1, 1, shape)
    X = orc.rfftn(x)
    assert orc.rel_l2(X, np.fft.rfftn(x)) < 2e-15
    # c2r of a Hermitian half-spectrum = unnormalized inverse
    y = orc.irfftn(X, shape[-1])
    assert orc.rel_l2(y / x.size, x) < 2e-15
    assert orc.rel_l2(y, np.fft.irfftn(X, s=shape) * x.size) < 2e-15


@pytest.mark.parametrize("n", [8, 9, 100])
def test_c2r_ignores_imag_of_dc_and_nyquist(orc, n):
    """1-D c2r reads only Re X[0] and (n even) Re X[n/2], as the hc2r
    codelets do (rdft/codelet-rdft.h:81-84 kr2c inputs)."""
    rng = np.random.default_rng(n)
    X = crand(rng, n // 2 + 1)
    y0 = orc.irfftn(X, n)
    X2 = X.copy()
    X2[0] += 5j
    if n % 2 == 0:
        X2[-1] -= 3j
    assert np.array_equal(orc.irfftn(X2, n), y0)
    full = np.zeros(n, complex)
    full[: n // 2 + 1] = X
    full[0] = X[0].real
    if n % 2 == 0:
        full[n // 2] = X[n // 2].real
    for k in range(1, (n + 1) // 2):
        full[n - k] = np.conj(X[k])
    assert orc.rel_l2(y0, (np.fft.ifft(full) * n).real) < 2e-15


def test_sampled_bins_large(orc):
    n = 1 << 16
    rng = np.random.default_rng(3)
    x = crand(rng, n)
    bins = np.array([0, 1, 777, n // 2, n - 1])
    assert orc.rel_l2(orc.dft_bins(x, bins), orc.fftn(x)[bins]) < 1e-15


def test_splitmix64_known_answers(orc):
    # Reference outputs of SplitMix64 (Steele, Lea, Flood 2014) from state 0:
    # the first two values of the sequence.
    assert orc.splitmix64(0) == 0xE220A8397B1DCDAF
    assert orc.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    u = orc.uniform(1000, 1)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.1
    assert np.array_equal(orc.uniform(10, 1), orc.uniform(10, 1))
    assert not np.array_equal(orc.uniform(10, 1), orc.uniform(10, 2))


def test_twiddle_symmetries(orc):
    for n in (8, 12, 4096, 1 << 24):
        assert orc.cexp(n // 4, n) == complex(0.0, -1.0)
        assert orc.cexp(n // 2, n) == complex(-1.0, 0.0)
        for m in (1, 3, n // 3, n - 5):
            w, v = orc.cexp(m, n), orc.cexp(n - m, n)
            assert abs(w * v - 1) < 4e-16
            assert abs(w - np.exp(-2j * np.pi * m / n)) < 1e-15
            assert orc.cexp(m, n, +1) == np.conj(w)


def test_planted_mistakes_are_caught(orc):
    """The closed-form and Hermitian checks above fail for a transform with
    the wrong sign or the wrong Hermitian half (planted on the CPU)."""
    n = 12
    k = np.arange(n)
    ej = np.zeros(n, complex)
    ej[5] = 1
    want = np.exp(-2j * np.pi * ((5 * k) % n) / n)
    wrong_sign = orc.fftn(ej, +1)
    assert np.max(np.abs(wrong_sign - want)) > 0.5
    X = orc.rfftn(np.linspace(-1, 1, n))
    bad = np.conj(X)
    assert orc.rel_l2(orc.irfftn(bad, n) / n, np.linspace(-1, 1, n)) > 0.1


@pytest.mark.parametrize("shape", [(6, 10, 14), (16, 9), (7, 5, 4, 3), (1000,)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_dft_bins_nd_against_numpy(orc, shape, sign):
    """Sampled multi-D bins (the full-size checks use them) against
    numpy.fft.fftn, complex and real inputs, negative bin indices."""
    rng = np.random.default_rng(len(shape) * 7 + sign)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    F = np.fft.fftn(x) if sign < 0 else np.fft.ifftn(x) * x.size
    bins = np.stack([rng.integers(-n, n, 6) for n in shape], axis=1)
    got = orc.dft_bins_nd(x, bins, sign)
    want = np.array([F[tuple(b)] for b in bins])
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(F))
    xr = rng.standard_normal(shape)
    Fr = np.fft.fftn(xr) if sign < 0 else np.fft.ifftn(xr) * xr.size
    got = orc.dft_bins_nd(xr, bins, sign)
    want = np.array([Fr[tuple(b)] for b in bins])
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(Fr))
