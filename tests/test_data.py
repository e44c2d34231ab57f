"""The input side (CPU): the climate and dengue file readers against files
written in the reference's layout (ClimateData.cpp:53-126,
DengueData.cpp:49-112), degree-heterogeneous community sizes against the
multinomial law of calcCumulativeComSize (network.cpp:116-164), and mosquito
shares reaching the t = 0 census."""
import numpy as np
import pytest
from scipy import stats

from arbo_gpu_fit_amd import configs
from arbo_gpu_fit_amd.data import community_sizes, mosquito_shares, read_climate_csv, read_dengue_csv
from arbo_gpu_fit_amd.engine import ArboError, Engine, SimSpec


def test_climate_file(tmp_path):
    """Columns by header name in any order, quotes removed, other columns
    ignored; every value read as a float."""
    rng = np.random.default_rng(1)
    T = rng.uniform(15, 35, 50).astype(np.float32)
    H = rng.uniform(0.3, 0.95, 50).astype(np.float32)
    P = rng.uniform(0, 40, 50).astype(np.float32)
    path = tmp_path / "climate_smooth.csv"
    lines = ['"date","P","extra","T","H"']
    for i in range(50):
        lines.append(f'"2010-01-{i:02d}",{float(P[i])!r},7,"{float(T[i])!r}",{float(H[i])!r}')
    path.write_text("\n".join(str(x) for x in lines) + "\n")
    cl = read_climate_csv(str(path), scale_eip=2.0)
    assert np.array_equal(cl.temperature, T) and np.array_equal(cl.humidity, H)
    assert np.array_equal(cl.precipitation, P) and cl.scale_eip == 2.0
    bad = tmp_path / "bad.csv"
    bad.write_text('"T","H"\n1,2\n')
    with pytest.raises(ArboError):
        read_climate_csv(str(bad))
    bad.write_text('"T","H","P"\n1,x,2\n')
    with pytest.raises(ArboError):
        read_climate_csv(str(bad))
    with pytest.raises(ArboError):
        read_climate_csv(str(tmp_path / "missing.csv"))


def test_dengue_file(tmp_path):
    inc = np.arange(260, dtype=np.float32) * 3.5
    path = tmp_path / "dengue.csv"
    path.write_text('"week","incidence"\n' + "".join(f"{i},{float(inc[i])!r}\n" for i in range(260)))
    assert np.array_equal(read_dengue_csv(str(path)), inc)
    path.write_text('"week","cases"\n1,2\n')
    with pytest.raises(ArboError):
        read_dengue_csv(str(path))


def test_crlf_files(tmp_path):
    """A file saved with CRLF endings reads as the LF one, the last column
    of the header included."""
    path = tmp_path / "climate_crlf.csv"
    path.write_bytes(b'"date","T","H","P"\r\n"d0",20.5,0.75,3\r\n"d1",21.25,0.5,0\r\n')
    cl = read_climate_csv(str(path))
    assert cl.temperature.tolist() == [20.5, 21.25]
    assert cl.humidity.tolist() == [0.75, 0.5] and cl.precipitation.tolist() == [3.0, 0.0]
    path = tmp_path / "dengue_crlf.csv"
    path.write_bytes(b'"week","incidence"\r\n1,4\r\n2,9\r\n')
    assert read_dengue_csv(str(path)).tolist() == [4.0, 9.0]


def test_community_sizes_reject_bad_totals():
    with pytest.raises(ArboError):
        community_sizes(4, 4, 15, 1.0, seed=1)          # fewer individuals than communities
    with pytest.raises(ArboError):
        community_sizes(4, 4, 1 << 32, 1.0, seed=1)     # counts are 32-bit


@pytest.mark.parametrize("hetero,nb", [(0.0, 4), (2.0, 4), (1.5, 8)])
def test_community_sizes_follow_degree_power(hetero, nb):
    """One individual per community, then a multinomial with p_c proportional
    to degree^hetero (degree 2/3/4 on a von Neumann lattice, 3/5/8 on a Moore
    one): chi-square against the expected counts."""
    W, H, total = 12, 9, 200_000
    out = community_sizes(W, H, total, hetero, seed=11, neighbourhood=nb)
    assert out.sum() == total and out.min() >= 1
    p = mosquito_shares(W, H, hetero, neighbourhood=nb).astype(np.float64)
    p /= p.sum()
    rest = total - W * H
    chi2 = (((out - 1) - rest * p) ** 2 / (rest * p)).sum()
    assert stats.chi2.sf(chi2, W * H - 1) > 1e-3
    if hetero == 0.0:
        assert np.allclose(p, 1.0 / (W * H))
    else:
        corner, interior = out[0], out[W + 1]
        assert interior > corner
    assert np.array_equal(out, community_sizes(W, H, total, hetero, seed=11, neighbourhood=nb))
    assert not np.array_equal(out, community_sizes(W, H, total, hetero, seed=12, neighbourhood=nb))
    with pytest.raises(ArboError):
        community_sizes(W, H, W * H - 1, hetero, seed=1)


def test_mosquito_shares_set_the_census(oracle_lib):
    """Each community's t = 0 mosquitoes are stochRound(m_ratio N share_c)
    (no seasonality); shares that do not sum to 1 are refused."""
    W, H = 5, 4
    humans = community_sizes(W, H, 8000, 1.0, seed=3)
    share = mosquito_shares(W, H, 2.0)
    p = configs.base_params(season_amp=0.0, seed_frac_h=0.0, seed_frac_m=0.0)
    eng = Engine(oracle_lib, threads=2)
    eng.init([SimSpec(W, H, humans, p, 5, mosq_share=share)], 10)
    mw = eng.state(0)["m_word"]
    cap = np.ceil(p.m_ratio * 8000.0 * share.astype(np.float64)).astype(np.int64)
    alive = np.add.reduceat(((mw & 3) != 3).astype(np.int64), np.concatenate([[0], np.cumsum(cap)[:-1]]))
    assert mw.size == cap.sum()
    assert np.all(np.abs(alive - p.m_ratio * 8000.0 * share) <= 1.0)
    with pytest.raises(ArboError):
        eng.init([SimSpec(W, H, humans, p, 5, mosq_share=share * 1.1)], 10)
    eng.close()
