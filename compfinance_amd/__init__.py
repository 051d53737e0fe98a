"""MI355X Monte Carlo pricing and AAD-risk engine - Python binding.

Thin ctypes layer over the C ABI of ``libcfamd.so`` (``include/cfamd.h``).
Names follow the reference library (asavine/CompFinance): models
``BlackScholes`` / ``Dupire``, products ``European`` / ``UOC`` /
``Europeans``, ``NumericalParam``, the simulation templates
``mcParallelSimul`` / ``mcParallelSimulAAD`` (mcBase.h:314-400, :564-753), the
store functions ``putBlackScholes`` ... ``putEuropeans`` (store.h:38-205) and
the entry points ``value``, ``AADriskOne``, ``AADriskAggregate``,
``dupireAADRisk`` (main.h:43-411).

Every computation runs in the HIP kernels of the library; there is no CPU
fallback.  Importing works without a GPU (labels, timelines and validation
are host code); simulations need one.
"""
import ctypes
import os
from collections import OrderedDict

import numpy as np

__all__ = [
    "FP64", "FP32", "AAD_CHECKPOINT", "AAD_RECOMPUTE",
    "BlackScholes", "Dupire", "European", "UOC", "Europeans", "Asian", "Book", "NumericalParam",
    "mcParallelSimul", "mcParallelSimulAAD", "SimulResults", "AADSimulResults",
    "putBlackScholes", "putDupire", "putEuropean", "putBarrier", "putEuropeans", "putAsian", "putBook",
    "value", "AADriskOne", "AADriskAggregate", "AADriskMulti", "dupireAADRisk", "clearStore",
    "mcParallelSimulAADMulti",
    "CfError", "lib", "last_kernel_stats", "LIB_PATH", "Server", "dupireCalib", "dupireCalibView", "dupireSuperbucket", "dupireSuperbucketBump",
    "sobol_points",
]

FP64, FP32 = 0, 1
AAD_CHECKPOINT, AAD_RECOMPUTE = 0, 1

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcfamd.so")

_c_double_p = ctypes.POINTER(ctypes.c_double)


class cf_model(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("spot", ctypes.c_double),
        ("vol", ctypes.c_double), ("rate", ctypes.c_double), ("div", ctypes.c_double),
        ("spot_measure", ctypes.c_int),
        ("n_spots", ctypes.c_int), ("n_times", ctypes.c_int),
        ("spots", _c_double_p), ("times", _c_double_p), ("vols", _c_double_p),
        ("max_dt", ctypes.c_double),
    ]


class cf_product(ctypes.Structure):
    pass


cf_product._fields_ = [
    ("kind", ctypes.c_int), ("strike", ctypes.c_double), ("exercise", ctypes.c_double),
    ("settlement", ctypes.c_double), ("barrier", ctypes.c_double), ("maturity", ctypes.c_double),
    ("monitor_freq", ctypes.c_double), ("smooth", ctypes.c_double), ("call_put", ctypes.c_int),
    ("n_options", ctypes.c_int), ("maturities", _c_double_p), ("strikes", _c_double_p),
    ("fixing_freq", ctypes.c_double), ("snap_freq", ctypes.c_double),
    ("n_parts", ctypes.c_int), ("parts", ctypes.POINTER(cf_product)),
]


class cf_num(ctypes.Structure):
    _fields_ = [
        ("parallel", ctypes.c_int), ("use_sobol", ctypes.c_int), ("num_path", ctypes.c_int),
        ("seed1", ctypes.c_int), ("seed2", ctypes.c_int), ("precision", ctypes.c_int),
        ("aad_mode", ctypes.c_int), ("first_path", ctypes.c_int),
    ]


class cf_kernel_stats(ctypes.Structure):
    _fields_ = [
        ("main_kernel_ms", ctypes.c_double), ("reduce_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
        ("grid_blocks", ctypes.c_int), ("block_threads", ctypes.c_int), ("lds_bytes", ctypes.c_int),
        ("path_steps", ctypes.c_int64), ("setup_ms", ctypes.c_double), ("host_ms", ctypes.c_double),
        ("lean_layout", ctypes.c_int), ("enqueue_ms", ctypes.c_double), ("wait_ms", ctypes.c_double),
    ]


class CfError(RuntimeError):
    """Error raised by the engine (the reference raises std::runtime_error)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_LIB = None


def lib():
    """The loaded ``libcfamd.so``; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m compfinance_amd.build` "
                "(the engine has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        sig = {
            "cf_num_payoffs": (ctypes.c_int, [P(cf_product)]),
            "cf_num_params": (ctypes.c_int, [P(cf_model)]),
            "cf_payoff_label": (ctypes.c_int, [P(cf_product), ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
            "cf_param_label": (ctypes.c_int, [P(cf_model), ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
            "cf_product_timeline": (ctypes.c_int, [P(cf_product), _c_double_p, ctypes.c_int]),
            "cf_sim_timeline": (ctypes.c_int, [P(cf_model), P(cf_product), _c_double_p, P(ctypes.c_int), ctypes.c_int]),
            # the simulation entry points take raw addresses (see _addr): one
            # integer per buffer is the cheapest argument to convert
            "cf_mc_simul": (ctypes.c_int, [P(cf_model), P(cf_product), P(cf_num)] + [ctypes.c_void_p] * 3),
            "cf_mc_simul_aad": (ctypes.c_int, [P(cf_model), P(cf_product), P(cf_num)] + [ctypes.c_void_p] * 6),
            "cf_aad_footprint": (ctypes.c_int64, [P(cf_model), P(cf_product), P(cf_num)]),
            "cf_last_error": (ctypes.c_char_p, []),
            "cf_last_kernel_stats": (ctypes.c_int, [P(cf_kernel_stats)]),
            "cf_version": (ctypes.c_char_p, []),
            "cf_device_synchronize": (ctypes.c_int, []),
            "cf_sobol_points": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
            "cf_put_black_scholes": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_char_p]),
            "cf_put_dupire": (ctypes.c_int, [ctypes.c_double, _c_double_p, ctypes.c_int, _c_double_p, ctypes.c_int,
                                             _c_double_p, ctypes.c_double, ctypes.c_char_p]),
            "cf_put_european": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_char_p]),
            "cf_put_barrier": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, ctypes.c_char_p]),
            "cf_put_europeans": (ctypes.c_int, [_c_double_p, _c_double_p, ctypes.c_int, ctypes.c_char_p]),
            "cf_value": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, P(cf_num), _c_double_p, ctypes.c_int]),
            "cf_aad_risk_one": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, P(cf_num), ctypes.c_char_p,
                                               _c_double_p, ctypes.c_int, _c_double_p, _c_double_p, ctypes.c_int]),
            "cf_aad_risk_aggregate": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, P(ctypes.c_char_p),
                                                     _c_double_p, ctypes.c_int, P(cf_num), _c_double_p, ctypes.c_int,
                                                     _c_double_p, _c_double_p, ctypes.c_int]),
            "cf_dupire_aad_risk": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, P(ctypes.c_char_p), _c_double_p,
                                                  ctypes.c_int, P(cf_num), _c_double_p, _c_double_p, _c_double_p,
                                                  ctypes.c_int]),
            "cf_store_clear": (None, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc < 0:
        raise CfError(rc, lib().cf_last_error().decode())
    return rc


def _dptr(a):
    return a.ctypes.data_as(_c_double_p) if a is not None else None


def _addr(a):
    """Address of a contiguous float64 array's data, or None."""
    return a.ctypes.data if a is not None else None


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


# ---------------------------------------------------------------- models


class _Described:
    """Keeps the ctypes descriptor of a model or product, and the payoff and
    parameter counts the engine reports for it, until a public attribute is
    assigned again.  The descriptor points at the object's own arrays, so
    changes made to them in place are seen without a rebuild."""

    def __setattr__(self, name, value):
        object.__setattr__(self, name, value)
        if not name.startswith("_"):
            self.__dict__.pop("_dcache", None)

    def _cache(self):
        c = self.__dict__.get("_dcache")
        if c is None:
            c = {"d": self.descriptor()}
            self.__dict__["_dcache"] = c
        return c


def _desc(obj):
    """Descriptor of a model or product, cached where the class allows."""
    return obj._cache()["d"] if isinstance(obj, _Described) else obj.descriptor()


def _count(obj, d, key, fn):
    if isinstance(obj, _Described):
        c = obj._cache()
        if key not in c:
            c[key] = _check(fn(ctypes.byref(d)))
        return c[key]
    return _check(fn(ctypes.byref(d)))


class BlackScholes(_Described):
    """BlackScholes(spot, vol, spotMeasure=False, rate=0, div=0) (mcMdlBS.h:72-93)."""

    def __init__(self, spot, vol, spotMeasure=False, rate=0.0, div=0.0):
        self.spot, self.vol, self.spotMeasure, self.rate, self.div = (
            float(spot), float(vol), bool(spotMeasure), float(rate), float(div))

    def descriptor(self):
        d = cf_model()
        d.kind, d.spot, d.vol, d.rate, d.div = 0, self.spot, self.vol, self.rate, self.div
        d.spot_measure = int(self.spotMeasure)
        return d

    def parameters(self):
        return [self.spot, self.vol, self.rate, self.div]

    def numParams(self):
        return 4

    def parameterLabels(self):
        return _param_labels(self)


class Dupire(_Described):
    """Dupire(spot, spots, times, vols, maxDt=0.25) (mcMdlDupire.h:76-111).

    vols[i][j] = local vol at spots[i], times[j] (spot major)."""

    def __init__(self, spot, spots, times, vols, maxDt=0.25):
        self.spot = float(spot)
        self.spots = _f64(spots)
        self.times = _f64(times)
        self.vols = _f64(vols).reshape(len(self.spots), len(self.times))
        self.maxDt = float(maxDt)

    def descriptor(self):
        # attributes may have been reassigned: coerce them to contiguous
        # float64 again (kept on the object, so later in-place changes are
        # seen) and check the grid's shape before the engine reads it
        spots, times = _f64(self.spots).ravel(), _f64(self.times).ravel()
        vols = _f64(self.vols)
        if vols.size != spots.size * times.size:
            raise ValueError("vols must hold len(spots) x len(times) values (spot major)")
        vols = vols.reshape(spots.size, times.size)
        for name, a in (("spots", spots), ("times", times), ("vols", vols)):
            object.__setattr__(self, name, a)
        d = cf_model()
        d.kind, d.spot = 1, float(self.spot)
        d.n_spots, d.n_times = spots.size, times.size
        self._keep = (spots, times, vols)
        d.spots, d.times, d.vols = (_dptr(a) for a in self._keep)
        d.max_dt = float(self.maxDt)
        return d

    def parameters(self):
        return [self.spot] + list(self.vols.ravel())

    def numParams(self):
        return 1 + self.vols.size

    def parameterLabels(self):
        return _param_labels(self)


# ---------------------------------------------------------------- products


class European(_Described):
    """European(strike, exerciseDate, settlementDate=exerciseDate) (mcPrd.h:44-86)."""

    def __init__(self, strike, exerciseDate, settlementDate=None):
        self.strike, self.exerciseDate = float(strike), float(exerciseDate)
        self.settlementDate = float(settlementDate) if settlementDate else self.exerciseDate

    def descriptor(self):
        d = cf_product()
        d.kind, d.strike, d.exercise, d.settlement = 0, self.strike, self.exerciseDate, self.settlementDate
        return d

    def payoffLabels(self):
        return _payoff_labels(self)

    def timeline(self):
        return _product_timeline(self)


class UOC(_Described):
    """UOC(strike, barrier, maturity, monitorFreq, smooth, callPut=False) (mcPrd.h:149-208).

    Non-positive smoothing is replaced by 1e-8 as putBarrier does (store.h:153)."""

    def __init__(self, strike, barrier, maturity, monitorFreq, smooth, callPut=False):
        self.strike, self.barrier, self.maturity = float(strike), float(barrier), float(maturity)
        self.monitorFreq, self.smooth, self.callPut = float(monitorFreq), float(smooth), bool(callPut)

    def descriptor(self):
        d = cf_product()
        d.kind, d.strike, d.barrier, d.maturity = 1, self.strike, self.barrier, self.maturity
        d.monitor_freq, d.smooth, d.call_put = self.monitorFreq, self.smooth, int(self.callPut)
        return d

    def payoffLabels(self):
        return _payoff_labels(self)

    def timeline(self):
        return _product_timeline(self)


class Europeans(_Described):
    """Europeans over (maturity, strike) pairs, maturity major (store.h:184-205, mcPrd.h:305-336)."""

    def __init__(self, maturities, strikes):
        self.maturities, self.strikes = _f64(maturities), _f64(strikes)
        if self.maturities.shape != self.strikes.shape:
            raise ValueError("maturities and strikes must have the same length")

    def descriptor(self):
        # attributes may have been reassigned: coerced and checked again
        mats, strikes = _f64(self.maturities).ravel(), _f64(self.strikes).ravel()
        if mats.shape != strikes.shape:
            raise ValueError("maturities and strikes must have the same length")
        object.__setattr__(self, "maturities", mats)
        object.__setattr__(self, "strikes", strikes)
        d = cf_product()
        d.kind, d.n_options = 2, mats.size
        self._keep = (mats, strikes)
        d.maturities, d.strikes = _dptr(mats), _dptr(strikes)
        return d

    def payoffLabels(self):
        return _payoff_labels(self)

    def timeline(self):
        return _product_timeline(self)


class Asian(_Described):
    """Arithmetic-average option (no reference class; Product<T> interface,
    mcBase.h:122-152): max(mean of the spot on the fixing dates - strike, 0),
    paid at maturity.  Fixings at fixingFreq, 2 fixingFreq, ... up to the
    maturity, each moved to the nearest date of the grid 0, snapFreq, ...,
    maturity when snapFreq > 0 (monthly fixings on a weekly simulation grid)."""

    def __init__(self, strike, maturity, fixingFreq, snapFreq=0.0, callPut=False):
        self.strike, self.maturity = float(strike), float(maturity)
        self.fixingFreq, self.snapFreq, self.callPut = float(fixingFreq), float(snapFreq), bool(callPut)

    def descriptor(self):
        d = cf_product()
        d.kind, d.strike, d.maturity, d.call_put = 3, self.strike, self.maturity, int(self.callPut)
        d.fixing_freq, d.snap_freq = self.fixingFreq, self.snapFreq
        return d

    def payoffLabels(self):
        return _payoff_labels(self)

    def timeline(self):
        return _product_timeline(self)


class Book:
    """Several products on the same paths (no reference class): payoffs
    concatenated in part order, labelled "<part index>:<part label>".  Parts:
    European (settled at exercise), UOC, Europeans, Asian."""

    def __init__(self, parts):
        self.parts = list(parts)
        if not self.parts:
            raise ValueError("empty book")

    def descriptor(self):
        d = cf_product()
        arr = (cf_product * len(self.parts))(*[p.descriptor() for p in self.parts])
        self._keep = arr
        d.kind, d.n_parts, d.parts = 4, len(self.parts), ctypes.cast(arr, ctypes.POINTER(cf_product))
        return d

    def payoffLabels(self):
        return _payoff_labels(self)

    def timeline(self):
        return _product_timeline(self)


class NumericalParam:
    """NumericalParam (main.h:33-40) plus precision, AAD variant and first path."""

    def __init__(self, parallel=True, useSobol=True, numPath=65536, seed1=12345, seed2=1234,
                 precision=FP64, aadMode=AAD_CHECKPOINT, firstPath=0):
        self.parallel, self.useSobol, self.numPath = parallel, useSobol, int(numPath)
        self.seed1, self.seed2 = int(seed1), int(seed2)
        self.precision, self.aadMode, self.firstPath = int(precision), int(aadMode), int(firstPath)

    def descriptor(self):
        n = cf_num()
        n.parallel, n.use_sobol, n.num_path = int(bool(self.parallel)), int(bool(self.useSobol)), self.numPath
        n.seed1, n.seed2, n.precision, n.aad_mode, n.first_path = (
            self.seed1, self.seed2, self.precision, self.aadMode, self.firstPath)
        return n


def _payoff_labels(prd):
    d = prd.descriptor()
    n = _check(lib().cf_num_payoffs(ctypes.byref(d)))
    out = []
    buf = ctypes.create_string_buffer(512)
    for k in range(n):
        _check(lib().cf_payoff_label(ctypes.byref(d), k, buf, 512))
        out.append(buf.value.decode())
    return out


def _param_labels(mdl):
    d = mdl.descriptor()
    n = _check(lib().cf_num_params(ctypes.byref(d)))
    out = []
    buf = ctypes.create_string_buffer(512)
    for k in range(n):
        _check(lib().cf_param_label(ctypes.byref(d), k, buf, 512))
        out.append(buf.value.decode())
    return out


def _product_timeline(prd):
    d = prd.descriptor()
    n = _check(lib().cf_product_timeline(ctypes.byref(d), None, 0))
    out = np.zeros(n)
    lib().cf_product_timeline(ctypes.byref(d), _dptr(out), n)
    return out


def simulation_timeline(mdl, prd):
    """Simulation timeline of ``mdl`` for ``prd`` and its product-date flags."""
    dm, dp = mdl.descriptor(), prd.descriptor()
    n = _check(lib().cf_sim_timeline(ctypes.byref(dm), ctypes.byref(dp), None, None, 0))
    t = np.zeros(n)
    c = np.zeros(n, dtype=np.int32)
    lib().cf_sim_timeline(ctypes.byref(dm), ctypes.byref(dp), _dptr(t),
                          c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n)
    return t, c.astype(bool)


def sobol_points(first_path, n_paths, dim, precision=FP64, layout=-1):
    """Sobol points of paths first_path .. first_path + n_paths - 1 (path p
    uses index p + 1: skipTo and next, sobol.h:77-151) in `dim` dimensions, as
    the path kernels draw them on the device: (words [n_paths][dim] uint32,
    normals [n_paths][dim] float64 or float32).  layout: 0 FULL (rows), 1
    LEAN, 2 FULL interleaved by four dimensions and read four at a time (the
    FP32 adjoint kernels' form), -1 FULL when it fits."""
    words = np.zeros((int(n_paths), int(dim)), dtype=np.uint32)
    normals = np.zeros((int(n_paths), int(dim)), dtype=np.float64 if precision == FP64 else np.float32)
    _check(lib().cf_sobol_points(int(first_path), int(n_paths), int(dim), int(precision), int(layout),
                                 words.ctypes.data, normals.ctypes.data))
    return words, normals


def synchronize():
    """Wait for the device (bench timing brackets)."""
    _check(lib().cf_device_synchronize())


def last_kernel_stats():
    """Device timings of the last simulation call on this thread."""
    s = cf_kernel_stats()
    _check(lib().cf_last_kernel_stats(ctypes.byref(s)))
    return {f: getattr(s, f) for f, _ in cf_kernel_stats._fields_}


# ---------------------------------------------------------------- simulation


class SimulResults:
    def __init__(self, means, std_errs, payoffs):
        self.means, self.std_errs, self.payoffs = means, std_errs, payoffs


class AADSimulResults:
    """payoffs[nPath][nPay], aggregated[nPath], risks[nParam] (mcBase.h:405-422),
    plus the means the engine reduces on the device."""

    def __init__(self, payoffs, aggregated, risks, payoff_means, agg_mean):
        self.payoffs, self.aggregated, self.risks = payoffs, aggregated, risks
        self.payoff_means, self.agg_mean = payoff_means, agg_mean


def _num(num, nPath=None, precision=None):
    if num is None:
        num = NumericalParam(numPath=nPath if nPath is not None else 65536)
    n = num.descriptor()
    if nPath is not None:
        n.num_path = int(nPath)
    if precision is not None:
        n.precision = int(precision)
    return n


def mcParallelSimul(prd, mdl, nPath=None, num=None, per_path=True, precision=None):
    """mcParallelSimul (mcBase.h:314-400): per-path payoff matrix (nPath x nPay)
    when ``per_path``, plus device-reduced means and standard errors."""
    dm, dp = _desc(mdl), _desc(prd)
    n = _num(num, nPath, precision)
    npay = _count(prd, dp, "npay", lib().cf_num_payoffs)
    # one buffer for the means and the standard errors
    buf = np.zeros(2 * npay)
    means, ses = buf[:npay], buf[npay:]
    pp = np.zeros((n.num_path, npay)) if per_path else None
    base = buf.ctypes.data
    _check(lib().cf_mc_simul(ctypes.byref(dm), ctypes.byref(dp), ctypes.byref(n), base, base + 8 * npay, _addr(pp)))
    return SimulResults(means, ses, pp)


def mcParallelSimulAAD(prd, mdl, nPath=None, weights=None, num=None, per_path=False, precision=None):
    """mcParallelSimulAAD (mcBase.h:564-753) with the linear aggregator
    sum_k weights[k] * payoff_k (default: the first payoff, mcBase.h:425)."""
    dm, dp = _desc(mdl), _desc(prd)
    n = _num(num, nPath, precision)
    npay = _count(prd, dp, "npay", lib().cf_num_payoffs)
    nparam = _count(mdl, dm, "nparam", lib().cf_num_params)
    # one buffer: weights, payoff means, aggregate mean, risks
    buf = np.zeros(2 * npay + 1 + nparam)
    w, means, risks = buf[:npay], buf[npay:2 * npay], buf[2 * npay + 1:]
    if weights is None:
        w[0] = 1.0
    else:
        wa = _f64(weights).ravel()
        if wa.size != npay:
            raise ValueError("weights must have one entry per payoff (%d), got %d" % (npay, wa.size))
        w[:] = wa
    pp = np.zeros((n.num_path, npay)) if per_path else None
    pa = np.zeros(n.num_path) if per_path else None
    base = buf.ctypes.data
    _check(lib().cf_mc_simul_aad(ctypes.byref(dm), ctypes.byref(dp), ctypes.byref(n), base, base + 8 * npay,
                                 base + 16 * npay, base + 8 * (2 * npay + 1), _addr(pp), _addr(pa)))
    return AADSimulResults(pp, pa, risks, means, float(buf[2 * npay]))


def aad_footprint(prd, mdl, num):
    """Device bytes one AAD call would take (checkpoint buffers dominate)."""
    dm, dp = mdl.descriptor(), prd.descriptor()
    n = num.descriptor()
    return _check(lib().cf_aad_footprint(ctypes.byref(dm), ctypes.byref(dp), ctypes.byref(n)))


# ---------------------------------------------------------------- store + entry points
# The store and the entry points run in the library (C++ mirror of store.h /
# main.h behind the store-level C ABI); the Python side keeps the objects only
# to label results.

_model_store = {}
_product_store = {}


def _id(x):
    return x.encode() if isinstance(x, str) else x


def putBlackScholes(spot, vol, qSpot, rate, div, store):
    """store.h:38-54"""
    _check(lib().cf_put_black_scholes(float(spot), float(vol), int(bool(qSpot)), float(rate), float(div), _id(store)))
    _model_store[store] = BlackScholes(spot, vol, qSpot, rate, div)


def putDupire(spot, spots, times, vols, maxDt, store):
    """store.h:56-73; vols spot major"""
    m = Dupire(spot, spots, times, vols, maxDt)
    v = np.ascontiguousarray(m.vols)
    _check(lib().cf_put_dupire(m.spot, _dptr(m.spots), len(m.spots), _dptr(m.times), len(m.times), _dptr(v),
                               m.maxDt, _id(store)))
    _model_store[store] = m


def putEuropean(strike, exerciseDate, settlementDate, store):
    """store.h:128-142"""
    _check(lib().cf_put_european(float(strike), float(exerciseDate), float(settlementDate or 0.0), _id(store)))
    _product_store[store] = European(strike, exerciseDate, settlementDate)


def putBarrier(strike, barrier, maturity, monitorFreq, smooth, callPut, store):
    """store.h:144-163"""
    _check(lib().cf_put_barrier(float(strike), float(barrier), float(maturity), float(monitorFreq), float(smooth),
                                int(bool(callPut)), _id(store)))
    _product_store[store] = UOC(strike, barrier, maturity, monitorFreq, smooth if smooth > 0 else 1e-8, callPut)


def putEuropeans(maturities, strikes, store):
    """store.h:184-205"""
    p = Europeans(maturities, strikes)
    _check(lib().cf_put_europeans(_dptr(p.maturities), _dptr(p.strikes), len(p.maturities), _id(store)))
    _product_store[store] = p


def putAsian(strike, maturity, fixingFreq, snapFreq, callPut, store):
    """Arithmetic-average option (no reference counterpart)."""
    L = lib()
    L.cf_put_asian.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_char_p]
    _check(L.cf_put_asian(float(strike), float(maturity), float(fixingFreq), float(snapFreq), int(bool(callPut)),
                          _id(store)))
    _product_store[store] = Asian(strike, maturity, fixingFreq, snapFreq, callPut)


def putBook(parts, store):
    """Book of stored products on shared paths (no reference counterpart)."""
    L = lib()
    L.cf_put_book.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_char_p]
    _check(L.cf_put_book(_label_array(parts), len(parts), _id(store)))
    _product_store[store] = Book([_product_store[p] for p in parts])


def clearStore():
    lib().cf_store_clear()
    _model_store.clear()
    _product_store.clear()


def _num_desc(num):
    return ctypes.byref(num.descriptor())


def _labels(productId):
    p = _product_store.get(productId)
    return p.payoffLabels() if p is not None else []


def value(modelId, productId, num):
    """main.h:80-96: {identifiers, values} of the mean payoffs."""
    n = num.descriptor()
    cap = 4096
    out = np.zeros(cap)
    npay = _check(lib().cf_value(_id(modelId), _id(productId), ctypes.byref(n), _dptr(out), cap))
    return OrderedDict(identifiers=_labels(productId), values=out[:npay].copy())


def _risk_result(modelId, productId, npar, pay, rv, risks):
    m = _model_store.get(modelId)
    return OrderedDict(payoffIds=_labels(productId), payoffValues=pay, riskPayoffValue=float(rv[0]),
                       paramIds=m.parameterLabels() if m is not None else [], risks=risks[:npar].copy())


def _caps(productId, modelId):
    p, m = _product_store.get(productId), _model_store.get(modelId)
    npay = len(p.payoffLabels()) if p is not None else 1
    npar = m.numParams() if m is not None else 1
    return npay, npar


def AADriskOne(modelId, productId, num, riskPayoff=""):
    """main.h:99-173"""
    npay, npar = _caps(productId, modelId)
    n = num.descriptor()
    pay, rv, risks = np.zeros(npay), np.zeros(1), np.zeros(npar)
    got = _check(lib().cf_aad_risk_one(_id(modelId), _id(productId), ctypes.byref(n),
                                       _id(riskPayoff) if riskPayoff else None, _dptr(pay), npay, _dptr(rv),
                                       _dptr(risks), npar))
    return _risk_result(modelId, productId, got, pay, rv, risks)


def _label_array(labels):
    arr = (ctypes.c_char_p * max(1, len(labels)))()
    for i, l in enumerate(labels):
        arr[i] = _id(l)
    return arr


def AADriskAggregate(modelId, productId, notionals, num):
    """main.h:176-254: notionals {payoff label: notional}."""
    npay, npar = _caps(productId, modelId)
    labs = list(notionals.keys())
    nots = _f64([notionals[k] for k in labs]) if labs else np.zeros(1)
    n = num.descriptor()
    pay, rv, risks = np.zeros(npay), np.zeros(1), np.zeros(npar)
    got = _check(lib().cf_aad_risk_aggregate(_id(modelId), _id(productId), _label_array(labs), _dptr(nots), len(labs),
                                             ctypes.byref(n), _dptr(pay), npay, _dptr(rv), _dptr(risks), npar))
    return _risk_result(modelId, productId, got, pay, rv, risks)


def dupireAADRisk(modelId, productId, notionals, num):
    """main.h:364-411: value, delta and the spots x times vega matrix."""
    m = _model_store.get(modelId)
    npar = m.numParams() if m is not None else 1
    labs = list(notionals.keys())
    nots = _f64([notionals[k] for k in labs]) if labs else np.zeros(1)
    n = num.descriptor()
    val, delta, vega = np.zeros(1), np.zeros(1), np.zeros(max(1, npar - 1))
    got = _check(lib().cf_dupire_aad_risk(_id(modelId), _id(productId), _label_array(labs), _dptr(nots), len(labs),
                                          ctypes.byref(n), _dptr(val), _dptr(delta), _dptr(vega), len(vega)))
    shape = (len(m.spots), len(m.times)) if isinstance(m, Dupire) else (got,)
    return OrderedDict(value=float(val[0]), delta=float(delta[0]), vega=vega[:got].reshape(shape))


# ---------------------------------------------------------------- request server
# Continuous batching of pricing and risk requests on one GPU (cf_server_*
# in cfamd.h); no counterpart in the reference.


class cf_server_opts(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("hbm_budget", ctypes.c_int64), ("auto_recompute", ctypes.c_int),
                ("graphs", ctypes.c_int), ("max_pack", ctypes.c_int)]


class cf_request(ctypes.Structure):
    _fields_ = [("model", cf_model), ("product", cf_product), ("num", cf_num), ("aad", ctypes.c_int),
                ("weights", _c_double_p)]


class cf_result(ctypes.Structure):
    _fields_ = [
        ("means", _c_double_p), ("std_errs", _c_double_p), ("cap_pay", ctypes.c_int),
        ("risks", _c_double_p), ("cap_risk", ctypes.c_int),
        ("status", ctypes.c_int), ("error", ctypes.c_char * 256), ("n_pay", ctypes.c_int), ("n_param", ctypes.c_int),
        ("agg_mean", ctypes.c_double), ("submit_ms", ctypes.c_double), ("queue_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double), ("device_bytes", ctypes.c_int64),
        ("recomputed", ctypes.c_int), ("packed", ctypes.c_int), ("plan_ms", ctypes.c_double),
        ("launch_ms", ctypes.c_double), ("device_ms", ctypes.c_double),
    ]


class cf_server_stats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in (
        "submitted", "completed", "failed", "queued", "running", "peak_running", "admitted_busy", "memory_waits",
        "recompute_switches", "arena_bytes", "peak_arena_bytes", "graph_captures", "graph_replays",
        "packed_launches", "packed_requests")]


def _server_lib():
    L = lib()
    if not getattr(L, "_server_sigs", False):
        P = ctypes.POINTER
        for name, res, args in [
            ("cf_server_create", ctypes.c_int, [P(cf_server_opts), P(ctypes.c_void_p)]),
            ("cf_server_submit", ctypes.c_int, [ctypes.c_void_p, P(cf_request), P(ctypes.c_int64)]),
            ("cf_server_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
            ("cf_server_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, P(cf_result)]),
            ("cf_server_get_stats", ctypes.c_int, [ctypes.c_void_p, P(cf_server_stats)]),
            ("cf_server_destroy", ctypes.c_int, [ctypes.c_void_p]),
        ]:
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        L._server_sigs = True
    return L


class Server:
    """Request server on the current GPU: `slots` requests on the device at
    once, device blocks from an arena of `hbmBudget` bytes; AAD requests that
    do not fit switch to the recompute variant when `autoRecompute`; with
    `graphs`, repeated requests replay a captured HIP graph of their launch
    sequence; queued requests of one path kernel too small to fill the device
    run up to `maxPack` to a launch (1: no packing)."""

    def __init__(self, slots=4, hbmBudget=16 << 30, autoRecompute=True, graphs=True, maxPack=8):
        L = _server_lib()
        o = cf_server_opts(int(slots), int(hbmBudget), int(bool(autoRecompute)), int(bool(graphs)), int(maxPack))
        h = ctypes.c_void_p()
        _check(L.cf_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._sizes = {}

    def submit(self, prd, mdl, num=None, aad=False, weights=None):
        """Queues one request; returns its id.  aad: prices, aggregate of
        sum_k weights[k] payoff_k (default: first payoff) and risks."""
        r = cf_request()
        r.model, r.product = mdl.descriptor(), prd.descriptor()
        r.num = (num or NumericalParam()).descriptor()
        r.aad = int(bool(aad))
        npay = _check(lib().cf_num_payoffs(ctypes.byref(r.product)))
        npar = _check(lib().cf_num_params(ctypes.byref(r.model))) if aad else 0
        w = None
        if aad and weights is not None:
            w = _f64(weights)
            if w.shape != (npay,):
                raise ValueError("weights must have one entry per payoff")
            r.weights = _dptr(w)
        rid = ctypes.c_int64()
        _check(_server_lib().cf_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        self._sizes[rid.value] = (npay, npar)
        return rid.value

    def poll(self, rid):
        return bool(_check(_server_lib().cf_server_poll(self._h, int(rid))))

    def wait(self, rid):
        npay, npar = self._sizes.pop(rid)
        means, ses, risks = np.zeros(npay), np.zeros(npay), np.zeros(max(npar, 1))
        out = cf_result()
        out.means, out.std_errs, out.cap_pay = _dptr(means), _dptr(ses), npay
        out.risks, out.cap_risk = _dptr(risks), npar
        rc = _server_lib().cf_server_wait(self._h, int(rid), ctypes.byref(out))
        if rc < 0:
            raise CfError(rc, out.error.decode() or lib().cf_last_error().decode())
        return OrderedDict(means=means, std_errs=ses, agg_mean=out.agg_mean, risks=risks[:npar],
                           submit_ms=out.submit_ms, queue_ms=out.queue_ms, latency_ms=out.latency_ms,
                           kernel_ms=out.kernel_ms, device_bytes=out.device_bytes, recomputed=bool(out.recomputed),
                           packed=out.packed, plan_ms=out.plan_ms, launch_ms=out.launch_ms,
                           device_ms=out.device_ms)

    def stats(self):
        s = cf_server_stats()
        _check(_server_lib().cf_server_get_stats(self._h, ctypes.byref(s)))
        return OrderedDict((f, getattr(s, f)) for f, _ in cf_server_stats._fields_)

    def close(self):
        if self._h:
            _check(_server_lib().cf_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------- local-vol calibration (FP64, host)


class cf_merton_ivs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("spot", "vol", "jmp_intens", "jmp_average", "jmp_std")]


def _calib_lib():
    L = lib()
    if not getattr(L, "_calib_sigs", False):
        P, d, i = ctypes.POINTER, _c_double_p, ctypes.c_int
        L.cf_dupire_calib.argtypes = [P(cf_merton_ivs), d, i, ctypes.c_double, d, i, ctypes.c_double, d, d, d,
                                      P(i), P(i), i]
        L.cf_superbucket_vega.argtypes = [P(cf_merton_ivs), d, i, ctypes.c_double, d, i, ctypes.c_double, d, i, d,
                                          i, d, i, d]
        L.cf_dupire_superbucket.argtypes = [P(cf_merton_ivs), ctypes.c_double, ctypes.c_char_p,
                                            P(ctypes.c_char_p), d, i, d, i, ctypes.c_double, d, i, ctypes.c_double,
                                            d, i, d, i, P(cf_num), d, d, d]
        L.cf_dupire_superbucket_bump.argtypes = L.cf_dupire_superbucket.argtypes
        L.cf_dupire_calib_view.argtypes = [P(cf_merton_ivs), d, i, ctypes.c_double, d, i, ctypes.c_double, d, i, d,
                                           i, d, d, d, d, P(i), P(i), i]
        L._calib_sigs = True
    return L


def dupireCalib(inclSpots, maxDs, inclTimes, maxDt, spot, vol, jmpIntens=0.0, jmpAverage=0.0, jmpStd=0.0):
    """dupireCalib (main.h:414-438): local vols on a Merton implied surface;
    returns {spots, times, lVols [spots x times]}."""
    L = _calib_lib()
    ivs = cf_merton_ivs(spot, vol, jmpIntens, jmpAverage, jmpStd)
    s, t = _f64(inclSpots), _f64(inclTimes)
    ns, nt = ctypes.c_int(), ctypes.c_int()
    args = (ctypes.byref(ivs), _dptr(s), len(s), float(maxDs), _dptr(t), len(t), float(maxDt))
    _check(L.cf_dupire_calib(*args, None, None, None, ctypes.byref(ns), ctypes.byref(nt), 0))
    S, T, V = np.zeros(ns.value), np.zeros(nt.value), np.zeros(ns.value * nt.value)
    _check(L.cf_dupire_calib(*args, _dptr(S), _dptr(T), _dptr(V), ctypes.byref(ns), ctypes.byref(nt), len(V)))
    return OrderedDict(spots=S, times=T, lVols=V.reshape(len(S), len(T)))


def dupireSuperbucket(spot, maxDt, productId, notionals, inclSpots, maxDs, inclTimes, maxDtVol, strikes, mats, vol,
                      jmpIntens, jmpAverage, jmpStd, num):
    """dupireSuperbucket (main.h:453-569): value, delta and the risks to the
    implied-vol spreads on strikes x mats, through the local-vol calibration
    and the GPU's local-vol risks."""
    L = _calib_lib()
    ivs = cf_merton_ivs(spot, vol, jmpIntens, jmpAverage, jmpStd)
    s, t, k, m = _f64(inclSpots), _f64(inclTimes), _f64(strikes), _f64(mats)
    labs = list(notionals.keys())
    nots = _f64([notionals[x] for x in labs]) if labs else np.zeros(1)
    val, delta, vega = np.zeros(1), np.zeros(1), np.zeros(len(k) * len(m))
    n = num.descriptor()
    _check(L.cf_dupire_superbucket(ctypes.byref(ivs), float(maxDt), _id(productId), _label_array(labs), _dptr(nots),
                                   len(labs), _dptr(s), len(s), float(maxDs), _dptr(t), len(t), float(maxDtVol),
                                   _dptr(k), len(k), _dptr(m), len(m), ctypes.byref(n), _dptr(val), _dptr(delta),
                                   _dptr(vega)))
    return OrderedDict(value=float(val[0]), delta=float(delta[0]), strikes=k, mats=m,
                       vega=vega.reshape(len(k), len(m)))


def dupireCalibView(inclSpots, maxDs, inclTimes, maxDt, spot, vol, jmpIntens, jmpAverage, jmpStd, strikes, mats,
                    spreads):
    """dupireCalib with a risk view (mcMdlDupire.h:337-388): local vols of the
    Merton surface plus implied-vol spreads [strikes x mats]."""
    L = _calib_lib()
    ivs = cf_merton_ivs(spot, vol, jmpIntens, jmpAverage, jmpStd)
    s, t, k, m = _f64(inclSpots), _f64(inclTimes), _f64(strikes), _f64(mats)
    sp = _f64(spreads).reshape(len(k), len(m))
    ns, nt = ctypes.c_int(), ctypes.c_int()
    args = (ctypes.byref(ivs), _dptr(s), len(s), float(maxDs), _dptr(t), len(t), float(maxDt), _dptr(k), len(k),
            _dptr(m), len(m), _dptr(sp))
    _check(L.cf_dupire_calib_view(*args, None, None, None, ctypes.byref(ns), ctypes.byref(nt), 0))
    S, T, V = np.zeros(ns.value), np.zeros(nt.value), np.zeros(ns.value * nt.value)
    _check(L.cf_dupire_calib_view(*args, _dptr(S), _dptr(T), _dptr(V), ctypes.byref(ns), ctypes.byref(nt), len(V)))
    return OrderedDict(spots=S, times=T, lVols=V.reshape(len(S), len(T)))


def dupireSuperbucketBump(spot, maxDt, productId, notionals, inclSpots, maxDs, inclTimes, maxDtVol, strikes, mats,
                          vol, jmpIntens, jmpAverage, jmpStd, num):
    """dupireSuperbucketBump (main.h:575-696): the superbucket by bump and
    reprice on the GPU (1e-8 spot bump, 1e-5 spread bumps, recalibrating the
    local vols for each)."""
    L = _calib_lib()
    ivs = cf_merton_ivs(spot, vol, jmpIntens, jmpAverage, jmpStd)
    s, t, k, m = _f64(inclSpots), _f64(inclTimes), _f64(strikes), _f64(mats)
    labs = list(notionals.keys())
    nots = _f64([notionals[x] for x in labs]) if labs else np.zeros(1)
    val, delta, vega = np.zeros(1), np.zeros(1), np.zeros(len(k) * len(m))
    n = num.descriptor()
    _check(L.cf_dupire_superbucket_bump(ctypes.byref(ivs), float(maxDt), _id(productId), _label_array(labs),
                                        _dptr(nots), len(labs), _dptr(s), len(s), float(maxDs), _dptr(t), len(t),
                                        float(maxDtVol), _dptr(k), len(k), _dptr(m), len(m), ctypes.byref(n),
                                        _dptr(val), _dptr(delta), _dptr(vega)))
    return OrderedDict(value=float(val[0]), delta=float(delta[0]), strikes=k, mats=m,
                       vega=vega.reshape(len(k), len(m)))


# ---------------------------------------------------------------- itemized risk


def mcParallelSimulAADMulti(prd, mdl, nPath=None, num=None, precision=None):
    """mcParallelSimulAADMulti (mcBase.h:858-989): (payoff means, risks
    [nParam][nPay]) with the risk of each payoff separately, same paths."""
    L = lib()
    L.cf_mc_simul_aad_multi.argtypes = [ctypes.POINTER(cf_model), ctypes.POINTER(cf_product), ctypes.POINTER(cf_num),
                                        _c_double_p, _c_double_p]
    dm, dp = mdl.descriptor(), prd.descriptor()
    n = _num(num, nPath, precision)
    npay = _check(L.cf_num_payoffs(ctypes.byref(dp)))
    npar = _check(L.cf_num_params(ctypes.byref(dm)))
    means, risks = np.zeros(npay), np.zeros((npar, npay))
    _check(L.cf_mc_simul_aad_multi(ctypes.byref(dm), ctypes.byref(dp), ctypes.byref(n), _dptr(means), _dptr(risks)))
    return means, risks


def AADriskMulti(modelId, productId, num):
    """main.h:269-312: {payoffs, params, values, risks [nParam][nPay]}."""
    L = lib()
    L.cf_aad_risk_multi.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(cf_num), _c_double_p,
                                    ctypes.c_int, _c_double_p, ctypes.c_int]
    npay, npar = _caps(productId, modelId)
    vals, risks = np.zeros(npay), np.zeros(npay * npar)
    n = num.descriptor()
    got = _check(L.cf_aad_risk_multi(_id(modelId), _id(productId), ctypes.byref(n), _dptr(vals), npay, _dptr(risks),
                                      len(risks)))
    m = _model_store.get(modelId)
    return OrderedDict(payoffs=_labels(productId), params=m.parameterLabels() if m is not None else None,
                       values=vals, risks=risks[:got * npay].reshape(got, npay))
