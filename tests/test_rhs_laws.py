"""The seeded right-hand-side recipes follow the laws the workloads state
(the GPU generators equal these fields, tests/test_gpu_parity.py):

  * normal:   i.i.d. N(0,1) values, mean-subtracted over the solver region;
  * gauss:    64 sources, centres uniform in [0,1]^3, widths uniform in
              [0.01, 0.1], amplitudes N(0,1), f = sum a exp(-|x-c|^2 / 2w^2);
  * spectrum: DST modes 1..K per direction with coefficients N(0,1) / |k|^2.

The random stream is the splitmix64 generator (pinned by its published
outputs); each law is checked on the generated values (Kolmogorov-Smirnov at
p > 1e-3, moments within 5 standard errors) and each field against its closed
form from the drawn parameters (numpy) or by inverting it (DST-I)."""
import ctypes

import numpy as np
import pytest
from scipy import fft, stats

from oracle import oracle as O


def lib():
    L = O.lib()
    L.og_mix64.argtypes = [ctypes.c_uint64]
    L.og_mix64.restype = ctypes.c_uint64
    L.og_rng_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
    L.og_rng_normal.restype = ctypes.c_double
    L.og_rhs_gauss_source.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double * 3, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
    L.og_rhs_spec_coef.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.og_rhs_spec_coef.restype = ctypes.c_double
    return L


def test_splitmix64_published_outputs():
    """SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c) seeded with
    1234567 yields 6457827717110365317, 3203168211198807973, ...: og_mix64(x) is
    the output for state x, the state advancing by 0x9E3779B97F4A7C15."""
    L = lib()
    expect = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
              16408922859458223821]
    s = 1234567
    for e in expect:
        assert L.og_mix64(s) == e
        s = (s + 0x9E3779B97F4A7C15) % (1 << 64)


def test_normal_field_law():
    n = (34, 34, 34)
    G = O.OracleGrid(n, tuple(1.0 / (v - 1) for v in n), (0, 0, 0), (0,) * 6)
    f = -G.rhs("normal", 2024)[G.solver_slice()].ravel()
    m = f.size
    assert abs(f.mean()) < 1e-14
    assert abs(f.var() - 1) < 5 * np.sqrt(2.0 / m)
    # the draws before the mean subtraction are N(0,1) (the mean moved them by ~1/sqrt(m))
    assert stats.kstest(f, "norm").pvalue > 1e-3
    rho = np.corrcoef(f[:-1], f[1:])[0, 1]  # neighbours along x
    assert abs(rho) < 5 / np.sqrt(m)


def test_gaussian_sources_law_and_field():
    L = lib()
    wmin, wmax = 0.01, 0.1
    c, w, a = (ctypes.c_double * 3)(), ctypes.c_double(), ctypes.c_double()
    P = []
    for seed in range(40):
        for s in range(64):
            L.og_rhs_gauss_source(seed, s, wmin, wmax, c, ctypes.byref(w), ctypes.byref(a))
            P.append((c[0], c[1], c[2], w.value, a.value))
    P = np.array(P)
    for d in range(3):
        assert P[:, d].min() > 0 and P[:, d].max() < 1
        assert stats.kstest(P[:, d], "uniform").pvalue > 1e-3
    assert P[:, 3].min() >= wmin and P[:, 3].max() <= wmax
    assert stats.kstest((P[:, 3] - wmin) / (wmax - wmin), "uniform").pvalue > 1e-3
    assert stats.kstest(P[:, 4], "norm").pvalue > 1e-3
    # the field of seed 3 is the closed-form sum over its 64 sources
    n = (26, 22, 18)
    ds = tuple(1.0 / (v - 1) for v in n)
    G = O.OracleGrid(n, ds, (0, 0, 0), (0,) * 6)
    b = G.rhs("gauss", 3, nsrc=64, wmin=wmin, wmax=wmax)
    src = P[3 * 64:4 * 64]
    Ls = G.lim_solver
    X = (np.arange(Ls[0], Ls[1]) - 1) * ds[0]
    Y = (np.arange(Ls[2], Ls[3]) - 1) * ds[1]
    Z = (np.arange(Ls[4], Ls[5]) - 1) * ds[2]
    Z, Y, X = np.meshgrid(Z, Y, X, indexing="ij")
    f = np.zeros_like(X)
    for cx, cy, cz, ws, am in src:
        f += am * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2 * ws * ws))
    assert np.abs(-b[G.solver_slice()] - f).max() <= 1e-13 * np.abs(f).max()


@pytest.mark.parametrize("K", [8, 16])
def test_spectrum_law_by_dst_inversion(K):
    """The field is sum_k c_k sin(kx pi i/(n-1)) ... on the n points of each
    direction, a DST-I series over the n-2 interior points: dstn recovers c_k
    times (n-1)^3, the modes above K vanish, and |k|^2 c_k is N(0,1)."""
    L = lib()
    n = 42
    G = O.OracleGrid((n,) * 3, (1.0 / (n - 1),) * 3, (0, 0, 0), (0,) * 6)
    seed = 9
    f = -G.rhs("spectrum", seed, K=K)[G.solver_slice()]
    coef = fft.dstn(f, type=1) / (n - 1) ** 3  # axes (z, y, x)
    k = np.arange(1, n - 1)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    inside = (kx <= K) & (ky <= K) & (kz <= K)
    assert np.abs(coef[~inside]).max() < 1e-12 * np.abs(coef[inside]).max()
    # against the drawn coefficients, mode by mode
    drawn = np.array([[[L.og_rhs_spec_coef(seed, K, x, y, z) for x in range(1, K + 1)] for y in range(1, K + 1)]
                      for z in range(1, K + 1)])
    assert np.abs(coef[:K, :K, :K] - drawn).max() < 1e-12 * np.abs(drawn).max()
    # the law: |k|^2 c_k ~ N(0,1)
    g = (coef * (kx ** 2 + ky ** 2 + kz ** 2))[inside]
    assert stats.kstest(g, "norm").pvalue > 1e-3
    assert abs(g.mean()) < 5 / np.sqrt(g.size) and abs(g.var() - 1) < 5 * np.sqrt(2.0 / g.size)
