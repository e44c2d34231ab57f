"""Pins the oracle's event-count laws (CPU) against what the mathematics fixes.

A source community with daily rate lambda sends its events three ways: some
stay at the source (noDistance), some go to a uniform random community
(longDistance), the rest are dispersed over its CSR row
(nepidemic.cu:64-136, mepidemic.cu:63-132).

* Stochastic rounding, the reference's rule (nepidemic.cu:67-70, 98-104):
  n = stochRound(lambda), stay = stochRound(n * noDistance),
  long-distance = stochRound((n - stay) * longDistance), the rest dispersed.
  Every count is the floor or the ceiling of its real value, exactly that
  value when it is an integer, and unbiased.
* Poisson: the three parts are independent Poisson variables with means
  lambda * noDistance, lambda (1 - noDistance) (1 - longDistance) and
  lambda (1 - noDistance) longDistance; with noDistance 0 the dispersed count
  is the same draw as before the stay split existed.
* Staying events are counted in the source's own column of its row; with the
  community not its own dispersal target, that column has weight 0.

A swapped fraction, a split applied to the wrong remainder or a lost
stochastic rounding moves a mean by many standard errors or breaks a
floor/ceiling bound.
"""
import ctypes as C
import math

import numpy as np
import pytest

from arbo_gpu_fit_amd import abi, configs
from arbo_gpu_fit_amd.engine import Engine, SimSpec

from .test_tables import table

NLOC = 20
F32 = np.float32


def _counts(lib, law, lam, omega, nd, n, seed=3, t=5, dir=1, v=2):
    local = np.zeros(n, np.int64)
    ld = np.zeros(n, np.int64)
    stay = np.zeros(n, np.int64)
    fx = C.c_uint64()
    st = C.c_uint32()
    for c in range(n):
        local[c] = lib.oracle_source_counts(law, lam, omega, nd, seed, c, t, dir, v, C.byref(fx), C.byref(st))
        ld[c] = fx.value
        stay[c] = st.value
    return local, ld, stay


def _floor_ceil(x, real):
    """Every x is the floor or the ceiling of its f32 real value."""
    real = np.asarray(real, np.float32).astype(np.float64)
    return np.all((x == np.floor(real)) | (x == np.ceil(real)))


@pytest.mark.parametrize("lam", [0.25, 3.0, 7.6, 155.3])
@pytest.mark.parametrize("nd", [0.0, 0.75])
def test_stoch_round_law(oracle_lib, lam, nd):
    n, omega = 20000, 0.2
    local, ld, stay = _counts(oracle_lib, abi.COUNT_STOCH_ROUND, lam, omega, nd, n)
    tot = local + ld + stay
    lam32 = float(F32(lam))
    assert _floor_ceil(tot, np.full(n, lam32))
    fr = lam32 - math.floor(lam32)
    assert abs(tot.mean() - lam32) <= 5 * math.sqrt(fr * (1 - fr) / n) + 1e-12
    # stay and long-distance: floor/ceiling of their f32 products, unbiased
    s_real = (tot.astype(np.float32) * F32(nd)).astype(np.float32)
    assert _floor_ceil(stay, s_real)
    assert abs((stay - s_real.astype(np.float64)).mean()) <= 5 * 0.5 / math.sqrt(n)
    l_real = ((tot - stay).astype(np.float32) * F32(omega)).astype(np.float32)
    assert _floor_ceil(ld, l_real)
    assert abs((ld - l_real.astype(np.float64)).mean()) <= 5 * 0.5 / math.sqrt(n)
    if nd == 0.0:
        assert np.all(stay == 0)


def test_stoch_round_integer_rates_are_exact(oracle_lib):
    local, ld, stay = _counts(oracle_lib, abi.COUNT_STOCH_ROUND, 12.0, 0.5, 0.5, 200)
    assert np.all(stay == 6) and np.all(ld == 3) and np.all(local == 3)


def test_poisson_law_without_stay_is_the_local_draw(oracle_lib):
    lam, omega = 37.5, 0.1
    local, fx, stay = _counts(oracle_lib, abi.COUNT_POISSON, lam, omega, 0.0, 300)
    mu_loc = float(F32(F32(lam) * (F32(1.0) - F32(omega))))
    tag = NLOC | (1 << 8) | (2 << 12)
    ref = np.array([oracle_lib.oracle_poisson(mu_loc, 3, c, 5, tag) for c in range(300)])
    assert np.array_equal(local, ref)
    assert np.all(stay == 0)
    assert np.all(fx == math.floor(float(F32(F32(lam) * F32(omega))) * 2.0 ** 24))


def test_poisson_law_with_stay(oracle_lib):
    lam, omega, nd, n = 20.0, 0.3, 0.4, 20000
    local, fx, stay = _counts(oracle_lib, abi.COUNT_POISSON, lam, omega, nd, n)
    for x, mu in ((stay, lam * nd), (local, lam * (1 - nd) * (1 - omega))):
        assert abs(x.mean() - mu) <= 5 * math.sqrt(mu / n)
        assert abs(x.var() - mu) <= 5 * math.sqrt((mu + 2 * mu * mu) / n)
    assert abs(np.corrcoef(stay, local)[0, 1]) < 5 / math.sqrt(n)
    go = F32(F32(lam) * (F32(1.0) - F32(nd)))
    assert np.all(fx == math.floor(float(F32(go * F32(omega))) * 2.0 ** 24))


def _cnt(lib, eng, C_):
    n = C.c_uint32()
    lib.oracle_tables(eng._h, 0, 13, None, 0, C.byref(n))
    cnt = np.zeros(n.value, np.uint32)
    lib.oracle_tables(eng._h, 0, 13, cnt.ctypes.data_as(C.POINTER(C.c_uint32)), n.value, C.byref(n))
    return cnt.reshape(C_, 2, 4, 9)


@pytest.mark.parametrize("law", [abi.COUNT_POISSON, abi.COUNT_STOCH_ROUND])
def test_staying_events_land_in_the_source_column(oracle_lib, law):
    """noDistance 1: every event stays; the source's own column carries them
    although the community is no dispersal target (von Neumann, no self)."""
    W = H = 3
    p = configs.reference_params(beta_hm=2.0, beta_mh=2.0, seed_frac_h=0.2, seed_frac_m=0.2, external_infection=0.0,
                                 no_distance=1.0, long_distance=0.0, count_law=law)
    eng = Engine(oracle_lib, threads=2)
    eng.init([SimSpec(W, H, np.full(9, 500, np.uint32), p, 8)], 1)
    eng.step(1)
    rp, to, thr = table(oracle_lib, eng, "row_ptr"), table(oracle_lib, eng, "to"), table(oracle_lib, eng, "thr")
    cnt = _cnt(oracle_lib, eng, 9)
    total = 0
    for c in range(9):
        row = to[rp[c]:rp[c + 1]]
        self_col = int(np.flatnonzero(row == c)[0])
        assert thr[rp[c] + self_col] == 0                 # weight 0: never a dispersal draw
        assert len(row) == 1 + sum(1 for d in range(9) if abs(d % 3 - c % 3) + abs(d // 3 - c // 3) == 1)
        other = np.delete(cnt[c], self_col, axis=-1)
        assert other.sum() == 0
        total += int(cnt[c][..., self_col].sum())
    assert total > 100
    eng.close()


def test_without_stay_split_the_row_has_no_self(oracle_lib):
    p = configs.reference_params(no_distance=0.0)
    eng = Engine(oracle_lib, threads=2)
    eng.init([SimSpec(3, 3, np.full(9, 50, np.uint32), p, 8)], 1)
    rp, to = table(oracle_lib, eng, "row_ptr"), table(oracle_lib, eng, "to")
    assert all(c not in to[rp[c]:rp[c + 1]] for c in range(9))
    eng.close()


def test_stay_fraction_of_dispersed_events(oracle_lib):
    """Over a lattice day, staying events / all row events ~ noDistance
    (long-distance 0): stochastic rounding keeps the ratio unbiased."""
    W = H = 4
    p = configs.reference_params(beta_hm=1.0, beta_mh=1.0, seed_frac_h=0.1, seed_frac_m=0.1, external_infection=0.0,
                                 no_distance=0.75, long_distance=0.0)
    assert p.count_law == abi.COUNT_STOCH_ROUND
    eng = Engine(oracle_lib, threads=2)
    eng.init([SimSpec(W, H, np.full(16, 1500, np.uint32), p, 9)], 1)
    eng.step(1)
    rp, to = table(oracle_lib, eng, "row_ptr"), table(oracle_lib, eng, "to")
    cnt = _cnt(oracle_lib, eng, 16)
    stay = row = 0
    for c in range(16):
        sc = int(np.flatnonzero(to[rp[c]:rp[c + 1]] == c)[0])
        stay += int(cnt[c][..., sc].sum())
        row += int(cnt[c].sum())
    assert row > 2000
    # per source the stay share is within 1/n of 0.75; summed, far tighter than this
    assert abs(stay / row - 0.75) < 0.01
    eng.close()


def test_count_law_validation(oracle_lib):
    for over in ({"count_law": 2}, {"no_distance": 1.5}, {"no_distance": -0.1}):
        eng = Engine(oracle_lib)
        with pytest.raises(RuntimeError):
            eng.init([SimSpec(2, 2, np.full(4, 10, np.uint32), configs.base_params(**over), 1)], 1)
        eng.close()
