"""Pins the oracle's host tables against closed forms (CPU).

Lifespan thresholds K[a] = floor(F(a) 2^24) against scipy's Weibull CDF and
the reference's bi-Weibull formula (ndemographic.cu:13-29, mdemographic.cu:13-17);
initial-age tables against the cumulative survival sums of
demographicintial.cu:48-83; the lattice CSR against Moore / von Neumann
geometry (network.cpp:31-41, ascending columns as network.cpp:95-113); alias
tables by exact reconstruction of each row's probabilities; the seasonal
density; Beta(2,5) immunity by a KS test against scipy.stats.beta.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats

from arbo_gpu_fit_amd import configs
from arbo_gpu_fit_amd.engine import Engine, SimSpec

ONE24 = 1 << 24
TABLES = {"death_h": 0, "death_m": 1, "init_h": 2, "init_m": 3, "row_ptr": 4, "to": 5, "thr": 6, "alias": 7,
          "in_ptr": 8, "in_src": 9, "in_slot": 10, "immune": 11, "cap": 12}


def table(lib, eng, name, sim=0):
    n = C.c_uint32()
    assert lib.oracle_tables(eng._h, sim, TABLES[name], None, 0, C.byref(n)) == 0
    out = np.zeros(n.value, np.uint32)
    assert lib.oracle_tables(eng._h, sim, TABLES[name], out.ctypes.data_as(C.POINTER(C.c_uint32)), n.value,
                             C.byref(n)) == 0
    return out


def make(lib, spec, T=10):
    eng = Engine(lib, threads=2)
    eng.init([spec], T)
    return eng


def f32(x):
    return float(np.float32(x))


def weibull_thresholds(n, scale, shape):
    a = np.arange(n, dtype=np.float64)
    F = stats.weibull_min.cdf(a, c=f32(shape), scale=f32(scale))
    return np.minimum(np.floor(F * ONE24), ONE24)


def check_close_table(got, ref, exact_frac=0.999):
    diff = np.abs(got.astype(np.int64) - ref.astype(np.int64))
    assert diff.max() <= 1, f"max diff {diff.max()} at {np.argmax(diff)}"
    assert (diff == 0).mean() >= exact_frac
    assert np.all(np.diff(got.astype(np.int64)) >= 0)


def test_lifespan_tables_weibull_exponential(oracle_lib):
    spec = SimSpec(2, 2, np.full(4, 10, np.uint32), configs.base_params(), 1)
    eng = make(oracle_lib, spec)
    p = spec.params
    Kh = table(oracle_lib, eng, "death_h")
    Km = table(oracle_lib, eng, "death_m")
    assert Kh.size == 65536 and Km.size == 2048
    check_close_table(Kh, weibull_thresholds(65536, p.h_scale, p.h_shape))
    check_close_table(Km, weibull_thresholds(2048, p.m_scale, p.m_shape))
    # lifetimes: E[min{a>=1: K[a] >= k}] over k uniform = sum_a P(age > a)
    mean_h = (1.0 - Kh[:-1] / ONE24).sum()
    assert abs(mean_h - 70 * 365 - 0.5) < 1.0
    mean_m = (1.0 - Km[:-1] / ONE24).sum()   # geometric with p = 1 - e^-0.1
    assert abs(mean_m - 1.0 / (1.0 - math.exp(-0.1))) < 1e-3
    eng.close()


def test_lifespan_tables_reference_biweibull(oracle_lib):
    p = configs.reference_params()
    spec = SimSpec(2, 2, np.full(4, 10, np.uint32), p, 1)
    eng = make(oracle_lib, spec)
    a = np.arange(65536, dtype=np.float64)
    haz = (a * f32(p.h_infant_scale)) ** f32(p.h_infant_shape)
    loc = f32(p.h_loc)
    haz = haz + np.where(a >= loc, (np.maximum(a - loc, 0) / f32(p.h_scale)) ** f32(p.h_shape), 0.0)
    ref = np.minimum(np.floor((1.0 - np.exp(-haz)) * ONE24), ONE24)
    check_close_table(table(oracle_lib, eng, "death_h"), ref)
    check_close_table(table(oracle_lib, eng, "death_m"), weibull_thresholds(2048, p.m_scale, p.m_shape))
    eng.close()


def test_initial_age_tables(oracle_lib):
    p = configs.base_params()
    eng = make(oracle_lib, SimSpec(2, 2, np.full(4, 10, np.uint32), p, 1))
    for name, n, step, scale, shape in (("init_h", 151, 365, p.h_scale, p.h_shape),
                                        ("init_m", 301, 1, p.m_scale, p.m_shape)):
        got = table(oracle_lib, eng, name)
        surv = stats.weibull_min.sf(np.arange(n) * step, c=f32(shape), scale=f32(scale))
        cum = np.cumsum(surv)
        ref = np.floor(cum / cum[-1] * ONE24)
        ref[-1] = ONE24
        check_close_table(got, ref, exact_frac=0.98)
    eng.close()


@pytest.mark.parametrize("nb,self_,sizes", [(8, 1, (4, 6, 9)), (4, 0, (2, 3, 4)), (8, 0, (3, 5, 8)),
                                            (4, 1, (3, 4, 5))])
def test_lattice_geometry(oracle_lib, nb, self_, sizes):
    W, H = 5, 4
    p = configs.base_params(neighbourhood=nb, include_self=self_)
    eng = make(oracle_lib, SimSpec(W, H, np.full(W * H, 3, np.uint32), p, 1))
    rp, to = table(oracle_lib, eng, "row_ptr"), table(oracle_lib, eng, "to")
    ip, isrc, islot = table(oracle_lib, eng, "in_ptr"), table(oracle_lib, eng, "in_src"), table(oracle_lib, eng, "in_slot")
    for c in range(W * H):
        x, y = c % W, c // W
        row = to[rp[c]:rp[c + 1]]
        edge = (x in (0, W - 1)) + (y in (0, H - 1))
        assert len(row) == {2: sizes[0], 1: sizes[1], 0: sizes[2]}[edge]
        assert np.all(np.diff(row.astype(np.int64)) > 0), "ascending columns"
        assert (c in row) == bool(self_)
        for d in row:
            dx, dy = abs(int(d) % W - x), abs(int(d) // W - y)
            assert max(dx, dy) <= 1 and (nb == 8 or dx + dy <= 1)
        # the in-edges of c are exactly the rows that contain c, ascending source
        srcs = isrc[ip[c]:ip[c + 1]]
        assert list(srcs) == sorted(s for s in range(W * H) if c in to[rp[s]:rp[s + 1]])
        for s, j in zip(srcs, islot[ip[c]:ip[c + 1]]):
            assert to[rp[s] + j] == c
    eng.close()


def test_isolated_community_keeps_events(oracle_lib):
    p = configs.base_params(include_self=0)
    eng = make(oracle_lib, SimSpec(1, 1, np.array([5], np.uint32), p, 1))
    assert list(table(oracle_lib, eng, "to")) == [0]
    eng.close()


@pytest.mark.parametrize("scale", [1.0, 0.0, 2.5])
def test_alias_tables_reconstruct_weights(oracle_lib, scale):
    W = H = 6
    p = configs.base_params(dispersal_scale=scale)
    eng = make(oracle_lib, SimSpec(W, H, np.full(W * H, 3, np.uint32), p, 1))
    rp, to = table(oracle_lib, eng, "row_ptr"), table(oracle_lib, eng, "to")
    thr, al = table(oracle_lib, eng, "thr"), table(oracle_lib, eng, "alias")
    for c in range(W * H):
        a, b = rp[c], rp[c + 1]
        n = b - a
        keep = thr[a:b] / ONE24
        prob = keep.copy()
        for j in range(n):
            prob[al[a + j]] += 1.0 - keep[j]
        prob /= n
        x, y = c % W, c // W
        d = np.array([math.hypot(int(t) % W - x, int(t) // W - y) for t in to[a:b]])
        w = np.exp(-d / scale) if scale > 0 else np.ones(n)
        assert np.abs(prob - w / w.sum()).max() <= n * 2.0 ** -24 + 1e-15
        if scale == 0:
            assert np.all(thr[a:b] == ONE24)
    eng.close()


def test_alias_destination_draws_chi_square(oracle_lib):
    """Destinations of many local events from one source follow the row weights."""
    p = configs.base_params(beta_hm=0.0, beta_mh=60.0, seed_frac_m=1.0, seed_frac_h=0.0, long_distance=0.0)
    W = H = 3
    eng = make(oracle_lib, SimSpec(W, H, np.full(9, 1000, np.uint32), p, 3), T=1)
    eng.step(1)
    n = C.c_uint32()
    oracle_lib.oracle_tables(eng._h, 0, 13, None, 0, C.byref(n))
    cnt = np.zeros(n.value, np.uint32)
    oracle_lib.oracle_tables(eng._h, 0, 13, cnt.ctypes.data_as(C.POINTER(C.c_uint32)), n.value, C.byref(n))
    cnt = cnt.reshape(9, 2, 4, 9)
    rp, to = table(oracle_lib, eng, "row_ptr"), table(oracle_lib, eng, "to")
    c = 4  # centre: 9 entries
    obs = cnt[c, 1].sum(axis=0)[:9].astype(float)
    d = np.array([math.hypot(int(t) % W - 1, int(t) // W - 1) for t in to[rp[c]:rp[c + 1]]])
    w = np.exp(-d)
    e = obs.sum() * w / w.sum()
    assert obs.sum() > 5000
    assert stats.chisquare(obs, e).pvalue > 1e-4
    eng.close()


def test_seasonal_density(oracle_lib):
    p = configs.base_params(season_amp=0.7, season_phase=1.3)
    eng = make(oracle_lib, SimSpec(2, 2, np.full(4, 10, np.uint32), p, 1), T=800)
    out = np.zeros(801, np.float32)
    oracle_lib.oracle_dens(eng._h, 0, out.ctypes.data_as(C.POINTER(C.c_float)), 801)
    t = np.arange(801)
    ref = (1.0 + f32(0.7) * np.sin(2 * np.pi * t / 365.0 + f32(1.3))).astype(np.float32)
    assert np.all(np.abs(out - ref) <= np.spacing(ref))
    eng.close()


def test_immunity_beta25(oracle_lib):
    p = configs.base_params(immunity_mode=1)
    eng = make(oracle_lib, SimSpec(60, 60, np.full(3600, 2, np.uint32), p, 4))
    f = table(oracle_lib, eng, "immune") / ONE24
    assert f.size == 3600 * 4
    assert stats.kstest(f, stats.beta(2, 5).cdf).pvalue > 1e-3
    eng.close()
    p0 = configs.base_params(immunity_mode=0, sero_prev=0.3)
    eng = make(oracle_lib, SimSpec(3, 3, np.full(9, 2, np.uint32), p0, 4))
    assert np.all(table(oracle_lib, eng, "immune") == math.floor(f32(0.3) * ONE24))
    eng.close()


def test_mosquito_capacity(oracle_lib):
    p = configs.base_params(season_amp=0.37)
    hum = np.array([1, 7, 200, 2000], np.uint32)
    eng = make(oracle_lib, SimSpec(2, 2, hum, p, 1))
    cap = table(oracle_lib, eng, "cap")
    assert list(cap) == [math.ceil(f32(2.0) * (1 + f32(0.37)) * int(n)) for n in hum]
    eng.close()
