"""ctypes mirror of include/arbo_mcmc.h and a thin Python driver over it.

`Mcmc` wraps one ArboMcmc handle: chains advanced by the C++ driver inside
libarbo_hip.so, their likelihoods from engine batches (or, for tests, from a
Python callback).  The history comes back as numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi
from .engine import ArboError, Engine

D = 7
SUR_INPUTS = 6   # surrogate inputs of parameters 0-4
PARAM_NAMES = ("beta_hm", "beta_mh", "log10_long_distance", "season_amp", "season_phase", "reporting", "dispersion")
CLIMATE_PARAM_NAMES = ("scale_eip", "scale_mle", "scale_cc", "scale_humidity", "min_mos_to_human", "reporting",
                       "dispersion")
# the reference's names of the climate set (MCMC.h:37-50)
REFERENCE_NAMES = ("scaleEIP", "scaleMLE", "scaleCC", "scaleHumidity", "minMosToHuman", "observationRate",
                   "dispersion")
PRIOR_UNIFORM, PRIOR_CIRCULAR, PRIOR_BETA, PRIOR_TRUNC_NORMAL, PRIOR_FIXED = 0, 1, 2, 3, 4
F_ZERO_WEEK_FLOOR = 1   # zero-mean weeks score the floor even for a zero count (the reference's fmax(-250, NaN))
TRUNC_HI = 10000.0   # upper end of the truncated-normal priors
FIT_SEASONAL, FIT_CLIMATE = 0, 1

_u32 = C.c_uint32
_f64p = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


class Prior(C.Structure):
    _fields_ = [("kind", _u32), ("reserved", _u32), ("a", C.c_double), ("b", C.c_double)]


class McmcDesc(C.Structure):
    _fields_ = [
        ("abi_version", _u32), ("n_chains", _u32), ("n_particles", _u32), ("n_weeks", _u32),
        ("skip_weeks", _u32), ("adapt_after", _u32), ("param_set", _u32), ("flags", _u32),
        ("n_blocks", _u32), ("reserved0", _u32),
        ("week_floor", C.c_double), ("mix_beta", C.c_double), ("init_scale", C.c_double),
        ("seed", C.c_uint64), ("priors", Prior * D),
    ]


class DaDesc(C.Structure):
    _fields_ = [("sub_steps", _u32), ("mode", _u32), ("step0", C.c_double), ("target_accept", C.c_double)]


DA_LANGEVIN = 0   # surrogate Langevin steps, then the simulator at the end point
DA_SCREEN = 1     # the adaptive-Metropolis proposal screened by the surrogate


class EngineApi(C.Structure):
    _fields_ = [("batch_init", C.c_void_p), ("batch_step", C.c_void_p), ("read_series", C.c_void_p),
                ("last_error", C.c_void_p)]


LOGLIK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, _u32, _f64p, _f64p)
McmcPtr = C.c_void_p

PROTOTYPES = {
    "arbo_mcmc_create": (C.c_int, [C.POINTER(McmcDesc), C.POINTER(abi.Shape), C.POINTER(abi.Params), _u32p, _f64p,
                                   C.POINTER(McmcPtr)]),
    "arbo_mcmc_run": (C.c_int, [McmcPtr, abi.EnginePtr, C.POINTER(EngineApi), _u32]),
    "arbo_mcmc_run_target": (C.c_int, [McmcPtr, LOGLIK_FN, C.c_void_p, _u32]),
    "arbo_mcmc_run_da": (C.c_int, [McmcPtr, abi.EnginePtr, C.POINTER(EngineApi), C.c_void_p, C.POINTER(DaDesc),
                                   _u32]),
    "arbo_mcmc_run_da_target": (C.c_int, [McmcPtr, LOGLIK_FN, C.c_void_p, C.c_void_p, C.POINTER(DaDesc), _u32]),
    "arbo_mcmc_features": (C.c_int, [McmcPtr, _u32, _f64p, C.POINTER(C.c_float)]),
    "arbo_mcmc_surrogate_logpost": (C.c_int, [McmcPtr, C.c_void_p, _u32, _f64p, _f64p, _f64p]),
    "arbo_mcmc_da_counters": (C.c_int, [McmcPtr, _u64p, _u64p, _f64p, _f64p]),
    "arbo_mcmc_iterations": (C.c_int, [McmcPtr, _u32p]),
    "arbo_mcmc_read": (C.c_int, [McmcPtr, _u32, _u32, _f64p, _f64p, C.POINTER(C.c_uint8)]),
    "arbo_mcmc_read_candidates": (C.c_int, [McmcPtr, _u32, _u32, _f64p, _f64p]),
    "arbo_mcmc_unsimulable": (C.c_int, [McmcPtr, _u64p]),
    "arbo_mcmc_counters": (C.c_int, [McmcPtr, _u64p, _u64p, _u64p, _u64p, _f64p]),
    "arbo_mcmc_sim_params": (C.c_int, [_f64p, C.POINTER(abi.Params), C.POINTER(abi.Params)]),
    "arbo_mcmc_sim_climate": (C.c_int, [_f64p, C.POINTER(abi.Climate), C.POINTER(abi.Climate)]),
    "arbo_mcmc_log_prior": (C.c_int, [McmcPtr, _f64p, _f64p]),
    "arbo_mcmc_destroy": (None, [McmcPtr]),
    "arbo_mcmc_last_error": (C.c_char_p, [McmcPtr]),
    "arbo_weekly_incidence": (C.c_int, [_u32p, _u32, _u32, _u32, _u32p]),
    "arbo_nb_loglik": (C.c_int, [_u32p, _u32p, _u32, _u32, C.c_double, C.c_double, C.c_double, _f64p]),
}


def bind(lib: C.CDLL) -> C.CDLL:
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def weekly_incidence(lib, incidence: np.ndarray, n_weeks: int) -> np.ndarray:
    """MCMC.cpp:199-209 over all serotypes (incidence [rows, S])."""
    inc = np.ascontiguousarray(incidence, dtype=np.uint32)
    rows, S = inc.shape
    out = np.zeros(n_weeks, np.uint32)
    rc = lib.arbo_weekly_incidence(inc.ctypes.data_as(_u32p), rows, S, n_weeks, out.ctypes.data_as(_u32p))
    if rc:
        raise ArboError(rc, "arbo_weekly_incidence")
    return out


def nb_loglik(lib, sim_weekly, data, skip: int, rho: float, r: float, floor: float = -250.0) -> float:
    sw = np.ascontiguousarray(sim_weekly, dtype=np.uint32)
    dt = np.ascontiguousarray(data, dtype=np.uint32)
    out = C.c_double()
    rc = lib.arbo_nb_loglik(sw.ctypes.data_as(_u32p), dt.ctypes.data_as(_u32p), sw.size, skip, rho, r, floor,
                            C.byref(out))
    if rc:
        raise ArboError(rc, "arbo_nb_loglik")
    return out.value


def engine_api(lib: C.CDLL) -> EngineApi:
    """The engine entry points of a library exporting include/arbo_c.h."""
    return EngineApi(*(C.cast(getattr(lib, n), C.c_void_p).value
                       for n in ("arbo_batch_init", "arbo_batch_step", "arbo_read_series", "arbo_last_error")))


@dataclass
class McmcConfig:
    """ArboMcmcDesc (include/arbo_mcmc.h).  n_particles: simulations per
    proposal (the reference's C_MAXSIMRUN is 30, constant.h:8); n_blocks: 0
    fresh particle seeds for every proposal, G >= 1 block pseudo-marginal
    (each proposal refreshes one of G blocks of the chain's seeds)."""
    n_chains: int
    n_weeks: int
    priors: list            # [(kind, a, b)] * 7
    n_particles: int = 1
    skip_weeks: int = 1
    adapt_after: int = 600
    week_floor: float = -250.0
    mix_beta: float = 0.05
    init_scale: float = 0.1
    seed: int = 1
    param_set: int = FIT_SEASONAL
    n_blocks: int = 0
    flags: int = 0

    def desc(self) -> McmcDesc:
        d = McmcDesc()
        d.abi_version = abi.ABI_VERSION
        d.n_chains = self.n_chains
        d.n_particles = self.n_particles
        d.n_weeks = self.n_weeks
        d.skip_weeks = self.skip_weeks
        d.adapt_after = self.adapt_after
        d.param_set = self.param_set
        d.flags = self.flags
        d.n_blocks = self.n_blocks
        d.week_floor = self.week_floor
        d.mix_beta = self.mix_beta
        d.init_scale = self.init_scale
        d.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        for i, (k, a, b) in enumerate(self.priors):
            d.priors[i] = Prior(k, 0, a, b)
        return d


class Mcmc:
    """Batched adaptive-Metropolis chains (include/arbo_mcmc.h)."""

    def __init__(self, cfg: McmcConfig, grid_w: int, grid_h: int, humans, base: abi.Params, data,
                 theta0=None, lib: C.CDLL | None = None, climate=None):
        """climate: an engine.ClimateSpec every simulation runs under (its
        coefficients are the parameters of the FIT_CLIMATE set); the driver
        copies it."""
        self.lib = bind(lib if lib is not None else abi.load_hip_library())
        self.cfg = cfg
        self._humans = np.ascontiguousarray(humans, dtype=np.uint32)
        self._data = np.ascontiguousarray(data, dtype=np.uint32)
        assert self._data.size == cfg.n_weeks
        shape = abi.Shape(grid_w, grid_h, self._humans.ctypes.data_as(_u32p), None)
        keep: list = []
        if climate is not None:
            cl = climate.struct(keep)
            keep.append(cl)
            shape.climate = C.pointer(cl)
        self.base = base
        t0 = None
        if theta0 is not None:
            self._theta0 = _f64(theta0).reshape(cfg.n_chains, D)
            t0 = self._theta0.ctypes.data_as(_f64p)
        self._h = McmcPtr()
        desc = cfg.desc()
        rc = self.lib.arbo_mcmc_create(C.byref(desc), C.byref(shape), C.byref(base),
                                       self._data.ctypes.data_as(_u32p), t0, C.byref(self._h))
        if rc:
            raise ArboError(rc, "arbo_mcmc_create")
        self._cb = None

    def _check(self, rc):
        if rc:
            msg = self.lib.arbo_mcmc_last_error(self._h)
            raise ArboError(rc, msg.decode() if msg else "")

    def run(self, engine: Engine, n_iters: int, api: EngineApi | None = None):
        """n_iters iterations with likelihoods from simulations on `engine`
        (whose library may differ from the driver's: pass engine_api(lib))."""
        self._check(self.lib.arbo_mcmc_run(self._h, engine._h, C.byref(api) if api is not None else None, n_iters))

    def run_target(self, loglik, n_iters: int):
        """n_iters iterations with loglik(theta[n, 7]) -> [n] from Python."""
        def cb(ctx, n, th, out):
            try:
                t = np.ctypeslib.as_array(th, shape=(n * D,)).reshape(n, D).copy()
                v = np.asarray(loglik(t), dtype=np.float64)
                for i in range(n):
                    out[i] = float(v[i])
                return 0
            except Exception:
                return -1
        self._cb = LOGLIK_FN(cb)
        self._check(self.lib.arbo_mcmc_run_target(self._h, self._cb, None, n_iters))

    def run_da(self, engine: Engine, sur, n_iters: int, sub_steps: int = 10, step0: float = 1.0,
               target_accept: float = 0.574, api: EngineApi | None = None, mode: int = DA_LANGEVIN):
        """n_iters delayed-acceptance iterations with the surrogate `sur` (a
        surrogate.Surrogate) as first stage: `sub_steps` Langevin steps on it
        (DA_LANGEVIN) or a screen of the adaptive-Metropolis proposal
        (DA_SCREEN); then the simulator."""
        da = DaDesc(sub_steps, mode, step0, target_accept)
        self._check(self.lib.arbo_mcmc_run_da(self._h, engine._h, C.byref(api) if api is not None else None, sur._h,
                                              C.byref(da), n_iters))

    def run_da_target(self, loglik, sur, n_iters: int, sub_steps: int = 10, step0: float = 1.0,
                      target_accept: float = 0.574, mode: int = DA_LANGEVIN):
        """Delayed acceptance with a Python second-stage log-likelihood."""
        def cb(ctx, n, th, out):
            try:
                t = np.ctypeslib.as_array(th, shape=(n * D,)).reshape(n, D).copy()
                v = np.asarray(loglik(t), dtype=np.float64)
                for i in range(n):
                    out[i] = float(v[i])
                return 0
            except Exception:
                return -1
        self._cb = LOGLIK_FN(cb)
        da = DaDesc(sub_steps, mode, step0, target_accept)
        self._check(self.lib.arbo_mcmc_run_da_target(self._h, self._cb, None, sur._h, C.byref(da), n_iters))

    def da_counters(self) -> dict:
        st, ac = C.c_uint64(), C.c_uint64()
        sec, step = C.c_double(), C.c_double()
        self._check(self.lib.arbo_mcmc_da_counters(self._h, C.byref(st), C.byref(ac), C.byref(sec), C.byref(step)))
        return {"surrogate_steps": st.value, "surrogate_accepted": ac.value, "surrogate_seconds": sec.value,
                "step": step.value}

    def features(self, theta) -> np.ndarray:
        t = _f64(np.atleast_2d(theta))
        x = np.zeros((t.shape[0], SUR_INPUTS), np.float32)
        self._check(self.lib.arbo_mcmc_features(self._h, t.shape[0], t.ctypes.data_as(_f64p),
                                                x.ctypes.data_as(C.POINTER(C.c_float))))
        return x

    def surrogate_logpost(self, sur, theta):
        t = _f64(np.atleast_2d(theta))
        lp = np.zeros(t.shape[0])
        g = np.zeros_like(t)
        self._check(self.lib.arbo_mcmc_surrogate_logpost(self._h, sur._h, t.shape[0], t.ctypes.data_as(_f64p),
                                                         lp.ctypes.data_as(_f64p), g.ctypes.data_as(_f64p)))
        return lp, g

    @property
    def iterations(self) -> int:
        n = C.c_uint32()
        self._check(self.lib.arbo_mcmc_iterations(self._h, C.byref(n)))
        return n.value

    def history(self, first: int = 0, count: int | None = None):
        """(theta [iters, chains, 7], loglik [iters, chains], accepted [iters, chains])."""
        n = self.iterations - first if count is None else count
        C_ = self.cfg.n_chains
        th = np.zeros((n, C_, D))
        ll = np.zeros((n, C_))
        acc = np.zeros((n, C_), np.uint8)
        self._check(self.lib.arbo_mcmc_read(self._h, first, n, th.ctypes.data_as(_f64p), ll.ctypes.data_as(_f64p),
                                            acc.ctypes.data_as(C.POINTER(C.c_uint8))))
        return th, ll, acc

    def candidates(self, first: int = 0, count: int | None = None):
        """(theta [iters, chains, 7], logpost [iters, chains]): each
        iteration's candidates and their log-likelihood plus log prior (-inf
        outside the support or unsimulable, NaN when a delayed-acceptance
        first stage stopped it)."""
        n = self.iterations - first if count is None else count
        C_ = self.cfg.n_chains
        th = np.zeros((n, C_, D))
        lp = np.zeros((n, C_))
        self._check(self.lib.arbo_mcmc_read_candidates(self._h, first, n, th.ctypes.data_as(_f64p),
                                                       lp.ctypes.data_as(_f64p)))
        return th, lp

    def counters(self) -> dict:
        p, s, sims, a = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        sec = C.c_double()
        self._check(self.lib.arbo_mcmc_counters(self._h, C.byref(p), C.byref(s), C.byref(sims), C.byref(a),
                                                C.byref(sec)))
        u = C.c_uint64()
        self._check(self.lib.arbo_mcmc_unsimulable(self._h, C.byref(u)))
        return {"proposals": p.value, "in_support": s.value, "simulations": sims.value, "accepted": a.value,
                "unsimulable": u.value, "eval_seconds": sec.value}

    def log_prior(self, theta) -> float:
        t = _f64(theta)
        out = C.c_double()
        rc = self.lib.arbo_mcmc_log_prior(self._h, t.ctypes.data_as(_f64p), C.byref(out))
        return out.value if rc == 0 else -np.inf

    def close(self):
        if self._h:
            self.lib.arbo_mcmc_destroy(self._h)
            self._h = McmcPtr()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sim_params(lib, theta, base: abi.Params) -> abi.Params:
    out = abi.Params()
    t = _f64(theta)
    rc = bind(lib).arbo_mcmc_sim_params(t.ctypes.data_as(_f64p), C.byref(base), C.byref(out))
    if rc:
        raise ArboError(rc, "arbo_mcmc_sim_params")
    return out


def sim_climate(lib, theta, base) -> abi.Climate:
    """The climate of one FIT_CLIMATE vector: base (abi.Climate) with its
    five coefficients set."""
    out = abi.Climate()
    t = _f64(theta)
    rc = bind(lib).arbo_mcmc_sim_climate(t.ctypes.data_as(_f64p), C.byref(base), C.byref(out))
    if rc:
        raise ArboError(rc, "arbo_mcmc_sim_climate")
    return out


PRIOR_NAMES = {PRIOR_UNIFORM: "Uniform", PRIOR_CIRCULAR: "CircularUniform", PRIOR_BETA: "Beta",
               PRIOR_TRUNC_NORMAL: "TruncatedNormal", PRIOR_FIXED: "Fixed"}


def _num(x: float) -> str:
    """A number as C++'s default ostream formatting prints it (6
    significant digits, %g), which the reference's writers use."""
    x = float(x)
    if np.isnan(x):
        return "nan" if not np.signbit(x) else "-nan"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    return f"{x:g}"


def write_chain_csv(path: str, theta: np.ndarray, logpost: np.ndarray, chain: int, names=PARAM_NAMES):
    """One chain's accepted states in the layout of MCMC::writeChain
    (MCMC.cpp:400-440): header "step,logLikelihood,<names>", then one line
    per iteration with the chain's current log posterior (the reference's
    prevAcceptedLogLikelihood includes the log prior, MCMC.cpp:284-300) and
    parameters.  theta [iterations, chains, 7] and logpost [iterations,
    chains] (history log-likelihood plus log prior)."""
    th = np.asarray(theta)[:, chain, :]
    lp = np.asarray(logpost)[:, chain]
    lines = ["step,logLikelihood," + ",".join(names)]
    for it in range(th.shape[0]):
        lines.append(f"{it},{_num(lp[it])}," + ",".join(_num(x) for x in th[it]))
    with open(path, "w") as f:
        f.write("\n".join(lines))


def write_candidates_csv(path: str, theta: np.ndarray, logpost: np.ndarray, accepted: np.ndarray, chain: int,
                         names=PARAM_NAMES):
    """One chain's candidates in the layout of MCMC::writeCandidates
    (MCMC.cpp:443-484): header "step,logLikelihood,<names>,accepted", then
    per iteration the candidate's log posterior (-inf outside the prior
    support, where the reference repeats its last value; nan when a
    delayed-acceptance first stage stopped it), its parameters and 1 if the
    chain accepted it.  theta and logpost as from Mcmc.candidates(),
    accepted from Mcmc.history()."""
    th = np.asarray(theta)[:, chain, :]
    lp = np.asarray(logpost)[:, chain]
    acc = np.asarray(accepted)[:, chain]
    lines = ["step,logLikelihood," + ",".join(names) + ",accepted"]
    for it in range(th.shape[0]):
        lines.append(f"{it},{_num(lp[it])}," + ",".join(_num(x) for x in th[it]) + f",{int(acc[it])}")
    with open(path, "w") as f:
        f.write("\n".join(lines))


# Columns of parameterData.csv (Parameter::writeParameters, parameter.cpp:8-47).
PARAMETER_DATA_COLUMNS = (
    "MCMCStep", "nSize", "minMosToHuman", "subPopTotal", "commSize", "bitingRate", "mnBitingSuccess",
    "nmBitingSuccess", "EIPRange", "EIP", "IIP", "Recovery", "noDistance", "EIRate", "longDistance", "initialSP",
    "mDemoScale", "mDemoShape", "nDemoScale", "nDemoShape", "nInfantScale", "nInfantShape", "nInfToDemoLoc",
    "netPower", "netM", "netSeed", "nHeteroPower", "mHeteroPower")


def parameter_data_row(step: int, params: abi.Params, humans, climate=None, biting_rate: float = 0.25) -> list:
    """One row of parameterData.csv for the simulations of a proposal.  The
    engine's parameters map onto the reference's SimulationParam fields:
    bitingRate x success = beta (nmBitingSuccess from beta_hm,
    mnBitingSuccess from beta_mh, at the climate's or the given biting
    rate), EIP = eip, IIP = the incubation midpoint, the lifespans, noDistance,
    externalInfection, longDistance, initialSeroPrev = sero_prev; minMosToHuman
    is the climate's or m_ratio.  Fields with no counterpart in this engine
    print the reference's defaults (SimulationParam.cpp:11-47): exIncPeriodRange
    2, netPower 1, netM 1, netSeed 1, nHeteroPopSize 0, mHeteroPopSize 0."""
    hum = np.asarray(humans)
    n = int(hum.sum())
    C_ = int(hum.size)
    br = float(climate.biting_rate) if climate is not None else biting_rate
    mmh = float(climate.min_mos_to_human) if climate is not None else float(params.m_ratio)
    return [step, n, mmh, C_, n / C_, br, params.beta_mh / br, params.beta_hm / br, 2.0, params.eip,
            0.5 * (params.iip_min + params.iip_max), params.recovery, params.no_distance,
            params.external_infection, params.long_distance, params.sero_prev, params.m_scale, params.m_shape,
            params.h_scale, params.h_shape, params.h_infant_scale, params.h_infant_shape, params.h_loc,
            1.0, 1.0, 1.0, 0.0, 0.0]


def write_parameter_data_csv(path: str, rows: list):
    """parameterData.csv in the layout of Parameter::writeParameters
    (parameter.cpp:8-47): the header, then one line per MCMC step whose
    candidate was simulated (rows from parameter_data_row)."""
    lines = [",".join(PARAMETER_DATA_COLUMNS)]
    for r in rows:
        lines.append(",".join([str(int(r[0]))] + [_num(x) for x in r[1:]]))
    with open(path, "w") as f:
        f.write("\n".join(lines))


def write_priors_csv(path: str, priors, names=PARAM_NAMES):
    """The priors in the layout of MCMC::writePriors (MCMC.cpp:383-397):
    "PARAMETER_NAME,PRIOR_DISTRIBUTION,PARAMETERS", the parameters as
    "(a;b)"."""
    lines = ["PARAMETER_NAME,PRIOR_DISTRIBUTION,PARAMETERS"]
    for name, (k, a, b) in zip(names, priors):
        lines.append(f"{name},{PRIOR_NAMES[k]},({a!r};{b!r})")
    with open(path, "w") as f:
        f.write("\n".join(lines))
