"""Pins the oracle's real-to-real transforms (orc_r2r) before the GPU path is
checked against them.

The oracle computes every kind as a complex DFT of its logical symmetric
sequence (oracle/oracle.c).  It is pinned here by:
  * brute-force sums of the definitions in doc/reference.texi:2021-2350
    (DCT-I..IV, DST-I..IV, DHT) and :913-930 (halfcomplex), written out
    term by term, for small n including the special cases n = 1 and 2;
  * scipy.fft.dct / dst types 1-4 with norm=None, whose documented
    definitions are these same unnormalized sums (an independent library:
    pocketfft's real transforms);
  * numpy.fft for R2HC / HC2R / DHT;
  * the inverse pairs and their logical sizes (doc/reference.texi:2147-2157
    and :2298-2309): REDFT00 twice gives 2(n-1) x, REDFT10 then REDFT01
    gives 2n x, REDFT11 twice 2n x, RODFT00 twice 2(n+1) x, DHT twice n x,
    R2HC then HC2R n x;
  * separability for multi-dimensional kinds and the guru stride layout.
A plausible mistake (a half-sample shift, the transposed type, the wrong
logical size, a dropped endpoint weight) fails the brute-force check;
test_planted_mistakes_are_caught shows it for the transposed type and the
endpoint weight of RODFT01.
"""
import numpy as np
import pytest
import scipy.fft as sf

R2HC, HC2R, DHT = 0, 1, 2
REDFT00, REDFT01, REDFT10, REDFT11 = 3, 4, 5, 6
RODFT00, RODFT01, RODFT10, RODFT11 = 7, 8, 9, 10
ALL = list(range(11))


def brute(kind, x):
    """Definitions of doc/reference.texi, summed term by term."""
    n = len(x)
    j = np.arange(n)
    y = np.zeros(n)
    for k in range(n):
        if kind == R2HC:
            if 2 * k <= n:
                y[k] = np.sum(x * np.cos(2 * np.pi * j * k / n))
            else:
                y[k] = -np.sum(x * np.sin(2 * np.pi * j * (n - k) / n))
        elif kind == HC2R:
            # Y_k = sum_j X_j exp(+2 pi i j k / n), X Hermitian from halfcomplex
            s = x[0]
            for m in range(1, n):
                if 2 * m < n:
                    re, im = x[m], x[n - m]
                elif 2 * m == n:
                    re, im = x[m], 0.0
                else:
                    re, im = x[n - m], -x[m]
                s += re * np.cos(2 * np.pi * m * k / n) - im * np.sin(2 * np.pi * m * k / n)
            y[k] = s
        elif kind == DHT:
            y[k] = np.sum(x * (np.cos(2 * np.pi * j * k / n) + np.sin(2 * np.pi * j * k / n)))
        elif kind == REDFT00:
            y[k] = x[0] + (-1) ** k * x[n - 1] + 2 * np.sum(
                x[1:n - 1] * np.cos(np.pi * j[1:n - 1] * k / (n - 1)))
        elif kind == REDFT10:
            y[k] = 2 * np.sum(x * np.cos(np.pi * (j + 0.5) * k / n))
        elif kind == REDFT01:
            y[k] = x[0] + 2 * np.sum(x[1:] * np.cos(np.pi * j[1:] * (k + 0.5) / n))
        elif kind == REDFT11:
            y[k] = 2 * np.sum(x * np.cos(np.pi * (j + 0.5) * (k + 0.5) / n))
        elif kind == RODFT00:
            y[k] = 2 * np.sum(x * np.sin(np.pi * (j + 1) * (k + 1) / (n + 1)))
        elif kind == RODFT10:
            y[k] = 2 * np.sum(x * np.sin(np.pi * (j + 0.5) * (k + 1) / n))
        elif kind == RODFT01:
            y[k] = (-1) ** k * x[n - 1] + 2 * np.sum(
                x[:n - 1] * np.sin(np.pi * (j[:n - 1] + 1) * (k + 0.5) / n))
        elif kind == RODFT11:
            y[k] = 2 * np.sum(x * np.sin(np.pi * (j + 0.5) * (k + 0.5) / n))
    return y


SCIPY = {REDFT00: (sf.dct, 1), REDFT10: (sf.dct, 2), REDFT01: (sf.dct, 3), REDFT11: (sf.dct, 4),
         RODFT00: (sf.dst, 1), RODFT10: (sf.dst, 2), RODFT01: (sf.dst, 3), RODFT11: (sf.dst, 4)}


def lib_ref(kind, x, axis=-1):
    """The same transform from an independent library along one axis."""
    n = x.shape[axis]
    if kind in SCIPY:
        f, t = SCIPY[kind]
        return f(x, type=t, axis=axis)
    X = np.fft.fft(x, axis=axis)
    if kind == DHT:
        return X.real - X.imag
    if kind == R2HC:
        X = np.moveaxis(X, axis, -1)
        k = np.arange(n)
        y = np.where(2 * k <= n, X.real, X[..., (n - k) % n].imag)
        return np.moveaxis(y, -1, axis)
    # HC2R: Hermitian extension then unnormalized inverse
    h = np.moveaxis(x, axis, -1)
    Z = np.zeros(h.shape, complex)
    Z[..., 0] = h[..., 0]
    for m in range(1, n):
        if 2 * m < n:
            Z[..., m] = h[..., m] + 1j * h[..., n - m]
        elif 2 * m == n:
            Z[..., m] = h[..., m]
        else:
            Z[..., m] = h[..., n - m] - 1j * h[..., m]
    return np.moveaxis(np.fft.ifft(Z, axis=-1).real * n, -1, axis)


def logical(kind, n):
    return {REDFT00: 2 * (n - 1), RODFT00: 2 * (n + 1)}.get(kind, n if kind <= DHT else 2 * n)


INVERSE = {R2HC: HC2R, HC2R: R2HC, DHT: DHT, REDFT00: REDFT00, REDFT10: REDFT01,
           REDFT01: REDFT10, REDFT11: REDFT11, RODFT00: RODFT00, RODFT10: RODFT01,
           RODFT01: RODFT10, RODFT11: RODFT11}


@pytest.mark.parametrize("kind", ALL)
def test_r2r_vs_bruteforce(orc, kind):
    rng = np.random.default_rng(10 + kind)
    for n in list(range(1, 21)) + [31, 64, 100]:
        if kind == REDFT00 and n == 1:
            continue
        x = rng.uniform(-1, 1, n)
        y = orc.r2rn(x, [kind])
        ref = brute(kind, x)
        assert np.max(np.abs(y - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), (kind, n)


@pytest.mark.parametrize("kind", ALL)
@pytest.mark.parametrize("n", [128, 1000, 1001, 4093, 4096])
def test_r2r_vs_library(orc, kind, n):
    rng = np.random.default_rng(n + kind)
    x = rng.uniform(-1, 1, (3, n))
    y = orc.r2rn(x, [kind], axes_batch=1)
    ref = lib_ref(kind, x, axis=-1)
    assert orc.rel_l2(y, ref) < 1e-14


@pytest.mark.parametrize("kind", ALL)
@pytest.mark.parametrize("n", [2, 3, 8, 15, 64, 257])
def test_inverse_pairs_and_logical_size(orc, kind, n):
    rng = np.random.default_rng(7 * n + kind)
    x = rng.uniform(-1, 1, n)
    y = orc.r2rn(orc.r2rn(x, [kind]), [INVERSE[kind]])
    assert orc.rel_l2(y / logical(kind, n), x) < 1e-15


def test_size_one_and_two(orc):
    # doc/reference.texi: REDFT00 is undefined for n = 1; size-1 R2HC / HC2R /
    # DHT / REDFT01 / RODFT01 are the identity (rdft/problem.c:129-133);
    # the others scale by 2 or sqrt(2)
    with pytest.raises(ValueError):
        orc.r2rn(np.ones(1), [REDFT00])
    x = np.array([0.75])
    want = {R2HC: 1, HC2R: 1, DHT: 1, REDFT01: 1, RODFT01: 1, REDFT10: 2, RODFT10: 2,
            RODFT00: 2, REDFT11: np.sqrt(2), RODFT11: np.sqrt(2)}
    for kind, f in want.items():
        assert abs(orc.r2rn(x, [kind])[0] - f * 0.75) < 1e-15
    # size 2: REDFT00, DHT and HC2R all equal R2HC (rdft/problem.c:194-197)
    x2 = np.array([0.3, -1.1])
    r = orc.r2rn(x2, [R2HC])
    for kind in (REDFT00, DHT, HC2R):
        assert np.allclose(orc.r2rn(x2, [kind]), r, atol=1e-15)


@pytest.mark.parametrize("kinds,shape", [((REDFT10, REDFT10), (16, 24)),
                                         ((REDFT01, RODFT11), (9, 14)),
                                         ((R2HC, DHT, RODFT00), (6, 5, 7)),
                                         ((REDFT00, HC2R, REDFT11), (5, 8, 3))])
def test_multid_separable(orc, kinds, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.uniform(-1, 1, (2,) + shape)
    y = orc.r2rn(x, kinds, axes_batch=1)
    ref = x
    for ax, k in enumerate(kinds):
        ref = lib_ref(k, ref, axis=ax + 1)
    assert orc.rel_l2(y, ref) < 1e-14


def test_guru_strides(orc):
    """Column-major input (transposed strides) and a padded output layout
    give the same numbers as the contiguous transform."""
    rng = np.random.default_rng(5)
    n0, n1 = 6, 10
    x = rng.uniform(-1, 1, (n0, n1))
    ref = orc.r2rn(x, [REDFT10, RODFT01])
    xt = np.ascontiguousarray(x.T).reshape(-1)
    out = np.full(n0 * (n1 + 3), np.nan)
    orc.guru_r2r([(n0, 1, n1 + 3), (n1, n0, 1)], [], [REDFT10, RODFT01], xt, out)
    y = out.reshape(n0, n1 + 3)[:, :n1]
    assert np.allclose(y, ref, rtol=0, atol=1e-14)
    assert np.all(np.isnan(out.reshape(n0, n1 + 3)[:, n1:]))


def test_planted_mistakes_are_caught(orc):
    """The brute-force check fails the transposed type (REDFT01 given for
    REDFT10) and a dropped endpoint weight (RODFT01 with x_{n-1} counted
    twice); both mistakes are planted on the CPU."""
    rng = np.random.default_rng(1)
    n = 9
    x = rng.uniform(-1, 1, n)
    x[n - 1] = 1.0
    assert np.max(np.abs(orc.r2rn(x, [REDFT01]) - brute(REDFT10, x))) > 0.1
    wrong = brute(RODFT01, x) + (-1) ** np.arange(n) * x[n - 1]
    assert np.max(np.abs(orc.r2rn(x, [RODFT01]) - wrong)) > 0.1
