"""Independent checks of oracle pieces that the other oracle tests leave open.

- The Gaussian generator (Box-Muller on two counter-based uniforms) against
  a restatement of SplitMix64 and the same transform in Python integer and
  float arithmetic, and against the normal distribution (moments, tail mass, a
  Kolmogorov-Smirnov bound).
- Strided, padded and batched r2c / c2r guru layouts against numpy.fft on
  the gathered lines (the layout semantics of doc/reference.texi:1062-1070
  and the r2c half-spectrum of doc/reference.texi:1923-1990).
- The committed golden fixtures against numpy.fft (pocketfft), an
  independent implementation, so that they pin more than the oracle's own
  past outputs.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def splitmix64(z):
    """SplitMix64 finalizer on Python ints (Steele, Lea, Flood 2014)."""
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def u01(seed, stream, i):
    return (splitmix64(seed ^ (stream << 40) ^ i) >> 11) / float(1 << 53)


def test_uniform_restated(orc):
    x = orc.uniform(64, 7, 3)
    want = np.array([2.0 * u01(7, 3, i) - 1.0 for i in range(64)])
    assert np.array_equal(x, want)


def test_gauss_restated(orc):
    count, seed, stream = 33, 4, 0
    x = orc.gauss(count, seed, stream)
    want = np.empty(count)
    for i in range(0, count, 2):
        r = math.sqrt(-2.0 * math.log(1.0 - u01(seed, stream, i)))
        th = 2.0 * math.pi * u01(seed, stream, i + 1)
        want[i] = r * math.cos(th)
        if i + 1 < count:
            want[i + 1] = r * math.sin(th)
    assert np.allclose(x, want, rtol=0, atol=1e-15)


def test_gauss_distribution(orc):
    n = 1 << 20
    x = orc.gauss(n, 4, 0)
    assert abs(x.mean()) < 5 / math.sqrt(n)
    assert abs(x.var() - 1.0) < 6 * math.sqrt(2.0 / n)
    # tail mass beyond 3 sigma: 0.0026998 of a standard normal
    tail = np.mean(np.abs(x) > 3.0)
    assert abs(tail - 0.0026998) < 5 * math.sqrt(0.0027 / n)
    # Kolmogorov-Smirnov distance against the normal CDF
    xs = np.sort(x[: 1 << 16])
    cdf = 0.5 * (1.0 + np.vectorize(math.erf)(xs / math.sqrt(2.0)))
    emp = np.arange(1, xs.size + 1) / xs.size
    d = max(np.max(emp - cdf), np.max(cdf - (emp - 1.0 / xs.size)))
    assert d < 1.63 / math.sqrt(xs.size)  # 1% critical value
    # even and odd elements (cos and sin halves) uncorrelated
    assert abs(np.corrcoef(x[0::2], x[1::2])[0, 1]) < 5 / math.sqrt(n / 2)


def test_r2c_strided_batched_layout(orc):
    """r2c of n = 12 reals at stride 3, howmany 4 at distance 50, written at
    stride 2 with distance 17 complex: against numpy on the gathered lines."""
    rng = np.random.default_rng(21)
    n, H, si, di, so, do = 12, 4, 3, 50, 2, 17
    nh = n // 2 + 1
    x = rng.uniform(-1, 1, 220)
    out = np.zeros(H * do + 20, complex)
    orc.guru_r2c([(n, si, so)], [(H, di, do)], x, out)
    for h in range(H):
        ref = np.fft.rfft(x[h * di + si * np.arange(n)])
        got = out[h * do + so * np.arange(nh)]
        assert orc.rel_l2(got, ref) < 2e-15
    # nothing outside the written positions
    mask = np.zeros(out.size, bool)
    for h in range(H):
        mask[h * do + so * np.arange(nh)] = True
    assert np.all(out[~mask] == 0)


def test_c2r_strided_batched_layout(orc):
    """c2r of n = 10 (6 bins at stride 2, distance 13) into reals at stride 3,
    distance 31: against numpy's inverse on the gathered bins, times n."""
    rng = np.random.default_rng(22)
    n, H, si, di, so, do = 10, 3, 2, 13, 3, 31
    nh = n // 2 + 1
    X = rng.uniform(-1, 1, 60) + 1j * rng.uniform(-1, 1, 60)
    out = np.zeros(H * do + 10)
    orc.guru_c2r([(n, si, so)], [(H, di, do)], X, out)
    for h in range(H):
        bins = X[h * di + si * np.arange(nh)].copy()
        # the c2r reads only Re of DC and Nyquist (as FFTW's r2cb codelets do)
        bins[0] = bins[0].real
        bins[-1] = bins[-1].real
        ref = np.fft.irfft(bins, n) * n
        assert orc.rel_l2(out[h * do + so * np.arange(n)], ref) < 2e-15


def test_r2c_2d_padded_rows(orc):
    """2-D r2c with padded rows on both sides (inembed / onembed)."""
    rng = np.random.default_rng(23)
    n0, n1, pin, pout = 6, 8, 11, 7
    nh = n1 // 2 + 1
    x = rng.uniform(-1, 1, n0 * pin)
    out = np.zeros(n0 * pout, complex)
    orc.guru_r2c([(n0, pin, pout), (n1, 1, 1)], [], x, out)
    ref = np.fft.rfft2(x.reshape(n0, pin)[:, :n1])
    assert orc.rel_l2(out.reshape(n0, pout)[:, :nh], ref) < 2e-15


def test_golden_fixtures_against_numpy():
    z = np.load(os.path.join(HERE, "golden", "golden.npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    import oracle
    for n, seed in meta["c2c"]:
        x = oracle.uniform_complex(n, seed)
        assert oracle.rel_l2(z["c2c_fwd_%d" % n], np.fft.fft(x)) < 5e-15, n
        assert oracle.rel_l2(z["c2c_bwd_%d" % n], np.fft.ifft(x) * n) < 5e-15, n
    for n, seed in meta["r2c"]:
        x = oracle.uniform(n, seed)
        assert oracle.rel_l2(z["r2c_%d" % n], np.fft.rfft(x)) < 5e-15, n
    checked = 0
    for key in z.files:
        if key.startswith("multi_"):
            shape = tuple(int(v) for v in key[len("multi_"):].split("x"))
            entry = [m for m in meta.get("multi", []) if tuple(m[0]) == shape]
            if entry:
                x = oracle.uniform_complex(int(np.prod(shape)), entry[0][1]).reshape(shape)
                assert oracle.rel_l2(z[key].reshape(shape), np.fft.fftn(x)) < 5e-15, key
                checked += 1
    assert checked == len(meta.get("multi", []))


def _r2r_numpy(kind, x):
    """FFTW's unnormalized r2r kinds (doc/reference.texi:2021-2350) from
    numpy / scipy, whose default (norm=None) DCT / DST types I-IV are the
    same sums."""
    import scipy.fft as sf
    n = x.size
    if kind == 0:  # R2HC: r_0..r_{n/2}, i_{(n+1)/2-1}..i_1
        X = np.fft.rfft(x)
        y = np.empty(n)
        y[: n // 2 + 1] = X.real
        for k in range(1, (n + 1) // 2):
            y[n - k] = X[k].imag
        return y
    if kind == 1:  # HC2R: the inverse of R2HC, unnormalized
        X = np.zeros(n // 2 + 1, complex)
        X.real = x[: n // 2 + 1]
        for k in range(1, (n + 1) // 2):
            X[k] += 1j * x[n - k]
        return np.fft.irfft(X, n) * n
    if kind == 2:  # DHT: sum x_j cas(2 pi j k / n) = Re X - Im X
        X = np.fft.fft(x)
        return X.real - X.imag
    typ = {3: 1, 4: 3, 5: 2, 6: 4, 7: 1, 8: 3, 9: 2, 10: 4}[kind]
    return sf.dct(x, type=typ) if kind <= 6 else sf.dst(x, type=typ)


def test_golden_c2r_and_r2r_fixtures_against_numpy():
    z = np.load(os.path.join(HERE, "golden", "golden.npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    import oracle
    for n, seed in meta["r2c"]:
        x = oracle.uniform(n, seed)
        # c2r(r2c(x)) = n x
        assert oracle.rel_l2(z["c2r_%d" % n], n * x) < 5e-15, n
    for n, seed in meta["r2r"]:
        x = oracle.uniform(n, seed)
        for kind in range(11):
            ref = _r2r_numpy(kind, x)
            assert oracle.rel_l2(z["r2r_%d_%d" % (kind, n)], ref) < 5e-14, (kind, n)
