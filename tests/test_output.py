"""The output files (CPU): arbo_write_ndata, arbo_write_nspatial and
arbo_write_nalive against restatements of the reference's writeIncidence
(writeoutput.cpp:13-98), writeSpatial (writeoutput.cpp:194-241) and
writeCensus (writeoutput.cpp:100-148), text for text, for a run that creates
the file and a run that appends to it."""
import numpy as np

from arbo_gpu_fit_amd.output import write_ndata


def reference_text(runs, S):
    out = "t,popSize" + "".join(f",incidence_DENV{v + 1}" for v in range(S))
    out += "".join(f",prevalence_DENV{v + 1}" for v in range(S))
    out += "".join(f",prevalence_DENV{v + 1}_random_community" for v in range(S)) + ",simRun"
    for r, ser in enumerate(runs):
        for t in range(ser["n_count"].size):
            out += f"\n{t},{ser['n_count'][t]}"
            out += "".join(f",{ser['n_incidence'][t, v]}" for v in range(S))
            out += "".join(f",{ser['n_infected'][t, v]}" for v in range(S))
            out += "".join(f",{ser['n_one_subpop'][t, v]}" for v in range(S))
            out += f",{r}"
    return out


def test_ndata_matches_reference_layout(tmp_path):
    rng = np.random.default_rng(1)
    S, rows = 3, 9
    runs = []
    for _ in range(2):
        runs.append({"n_count": rng.integers(0, 10**6, rows), "m_count": rng.integers(0, 10**6, rows),
                     "n_infected": rng.integers(0, 5000, (rows, S)), "n_incidence": rng.integers(0, 500, (rows, S)),
                     "n_one_subpop": rng.integers(0, 50, (rows, S))})
    path = str(tmp_path / "nData0.csv")
    for r, ser in enumerate(runs):
        write_ndata(path, ser, r)
    assert open(path).read() == reference_text(runs, S)


def reference_spatial_text(runs):
    """writeSpatial (writeoutput.cpp:194-241) restated: the simRun column only
    after the last community of each day."""
    S = runs[0].shape[1]
    out = "t,subPop" + "".join(f",prevalence.DENV{v + 1}" for v in range(S)) + ",simRun"
    for r, sp in enumerate(runs):
        rows, S, C = sp.shape
        for t in range(rows):
            for c in range(C):
                out += f"\n{t},{c}" + "".join(f",{sp[t, v, c]}" for v in range(S))
            out += f",{r}"
    return out


def test_nspatial_matches_reference_layout(tmp_path):
    from arbo_gpu_fit_amd.output import write_nspatial
    rng = np.random.default_rng(2)
    runs = [rng.integers(0, 3000, (6, 4, 5)), rng.integers(0, 3000, (6, 4, 5))]
    path = str(tmp_path / "nSpatial0.csv")
    for r, sp in enumerate(runs):
        write_nspatial(path, sp, r)
    assert open(path).read() == reference_spatial_text(runs)


def reference_alive_text(runs):
    """writeCensus (writeoutput.cpp:100-148) restated: the first
    min(C_NSIZERECORD = 5000, n) humans, each history column read at
    [s * n + i]."""
    S = runs[0][1].shape[0]
    out = "age" + "".join(f",history{v + 1}" for v in range(S)) + ",simRun"
    for r, (age, hist) in enumerate(runs):
        n = age.size
        for i in range(min(5000, n)):
            out += f"\n{age[i]}" + "".join(f",{hist[v, i]}" for v in range(S)) + f",{r}"
    return out


def test_nalive_matches_reference_layout(tmp_path):
    from arbo_gpu_fit_amd.output import write_nalive
    rng = np.random.default_rng(3)
    runs = []
    for n in (7000, 40):   # more humans than a census records, then fewer
        age = rng.integers(0, 36500, n)
        hist = rng.integers(-2, 30000, (4, n))
        runs.append((age, hist))
    path = str(tmp_path / "nAlive_0.csv")
    for r, (age, hist) in enumerate(runs):
        write_nalive(path, age, hist, r)
    text = open(path).read()
    assert text == reference_alive_text(runs)
    assert text.count("\n") == 5000 + 40


def test_naoi_matches_reference_layout(tmp_path):
    """writeAOI (writeoutput.cpp:152-191) restated: 5000 lines of the four
    rings side by side."""
    from arbo_gpu_fit_amd.output import write_naoi
    rng = np.random.default_rng(4)
    runs = [rng.integers(0, 30000, (4, 5000)), rng.integers(0, 30000, (4, 5000))]
    path = str(tmp_path / "nAOI_0.csv")
    for r, a in enumerate(runs):
        write_naoi(path, a, r)
    want = "first,second,third,fourth,simRun"
    for r, a in enumerate(runs):
        for i in range(5000):
            want += f"\n{a[0, i]},{a[1, i]},{a[2, i]},{a[3, i]},{r}"
    assert open(path).read() == want
