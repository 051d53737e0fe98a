"""ctypes binding of the CPU restatement (liboracle.so) and of the reference
driver (_ref/ref_driver).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry and by the
CPU-baseline leg of bench.py, as the checker.  Model and product arguments
are the same keyword dictionaries the engine's classes take.
"""
import ctypes
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

_dp = ctypes.POINTER(ctypes.c_double)


class or_model(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("spot", ctypes.c_double), ("vol", ctypes.c_double),
        ("rate", ctypes.c_double), ("div", ctypes.c_double), ("spot_measure", ctypes.c_int),
        ("n_spots", ctypes.c_int), ("n_times", ctypes.c_int),
        ("spots", _dp), ("times", _dp), ("vols", _dp), ("max_dt", ctypes.c_double),
    ]


class or_product(ctypes.Structure):
    pass


or_product._fields_ = [
    ("kind", ctypes.c_int), ("strike", ctypes.c_double), ("exercise", ctypes.c_double),
    ("settlement", ctypes.c_double), ("barrier", ctypes.c_double), ("maturity", ctypes.c_double),
    ("monitor_freq", ctypes.c_double), ("smooth", ctypes.c_double), ("call_put", ctypes.c_int),
    ("n_options", ctypes.c_int), ("maturities", _dp), ("strikes", _dp),
    ("fixing_freq", ctypes.c_double), ("snap_freq", ctypes.c_double),
    ("n_parts", ctypes.c_int), ("parts", ctypes.POINTER(or_product)),
]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(LIB_PATH)
        L.or_num_payoffs.argtypes = [ctypes.POINTER(or_product)]
        L.or_num_params.argtypes = [ctypes.POINTER(or_model)]
        L.or_timeline.argtypes = [ctypes.POINTER(or_model), ctypes.POINTER(or_product), _dp,
                                  ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.or_fill_data.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, _dp,
                                   ctypes.c_int]
        L.or_sobol_dirs.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint)]
        L.or_sobol_states.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.POINTER(ctypes.c_uint)]
        L.or_inv_normal.argtypes = [ctypes.c_long, _dp, _dp]
        L.or_value_rng.argtypes = [ctypes.POINTER(or_model), ctypes.POINTER(or_product), ctypes.c_long,
                                   ctypes.c_long, ctypes.c_int, ctypes.c_uint, ctypes.c_uint, _dp]
        L.or_aad_rng.argtypes = [ctypes.POINTER(or_model), ctypes.POINTER(or_product), ctypes.c_long, ctypes.c_int,
                                 _dp, ctypes.c_int, ctypes.c_uint, ctypes.c_uint, _dp, _dp, _dp]
        L.or_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


class Model:
    """Black-Scholes: Model(spot=, vol=, spotMeasure=, rate=, div=);
    Dupire: Model(spot=, spots=, times=, vols=, maxDt=)."""

    def __init__(self, **kw):
        self.kw = kw
        d = or_model()
        d.spot = kw["spot"]
        if "spots" in kw:
            d.kind = 1
            self._keep = (_arr(kw["spots"]), _arr(kw["times"]), _arr(kw["vols"]))
            d.n_spots, d.n_times = len(self._keep[0]), len(self._keep[1])
            d.spots, d.times, d.vols = (a.ctypes.data_as(_dp) for a in self._keep)
            d.max_dt = kw.get("maxDt", 0.25)
        else:
            d.kind = 0
            d.vol, d.rate, d.div = kw["vol"], kw.get("rate", 0.0), kw.get("div", 0.0)
            d.spot_measure = int(bool(kw.get("spotMeasure", False)))
        self.d = d

    def spec(self):
        kw = self.kw
        if self.d.kind == 1:
            s = ["model dupire", repr(float(kw["spot"])), repr(float(kw.get("maxDt", 0.25))), str(len(self._keep[0]))]
            s += [repr(float(x)) for x in self._keep[0]] + [str(len(self._keep[1]))] + [repr(float(x)) for x in self._keep[1]]
            s += [repr(float(x)) for x in self._keep[2].ravel()]
        else:
            s = ["model bs", repr(kw["spot"]), repr(kw["vol"]), str(int(bool(kw.get("spotMeasure", False)))),
                 repr(kw.get("rate", 0.0)), repr(kw.get("div", 0.0))]
        return " ".join(s)


class Product:
    """European: Product(strike=, exerciseDate=, settlementDate=);
    barrier: Product(strike=, barrier=, maturity=, monitorFreq=, smooth=, callPut=);
    Europeans: Product(maturities=, strikes=);
    Asian: Product(strike=, maturity=, fixingFreq=, snapFreq=, callPut=);
    book: Product(parts=[dict, ...])."""

    def __init__(self, **kw):
        self.kw = kw
        d = or_product()
        if "parts" in kw:
            d.kind = 4
            self._parts = [Product(**q) for q in kw["parts"]]
            self._arr = (or_product * len(self._parts))(*[q.d for q in self._parts])
            d.n_parts, d.parts = len(self._parts), ctypes.cast(self._arr, ctypes.POINTER(or_product))
        elif "fixingFreq" in kw:
            d.kind = 3
            d.strike, d.maturity = kw["strike"], kw["maturity"]
            d.fixing_freq, d.snap_freq = kw["fixingFreq"], kw.get("snapFreq", 0.0)
            d.call_put = int(bool(kw.get("callPut", False)))
        elif "barrier" in kw:
            d.kind = 1
            d.strike, d.barrier, d.maturity = kw["strike"], kw["barrier"], kw["maturity"]
            d.monitor_freq, d.smooth = kw["monitorFreq"], kw["smooth"]
            d.call_put = int(bool(kw.get("callPut", False)))
        elif "maturities" in kw:
            d.kind = 2
            self._keep = (_arr(kw["maturities"]), _arr(kw["strikes"]))
            d.n_options = len(self._keep[0])
            d.maturities, d.strikes = (a.ctypes.data_as(_dp) for a in self._keep)
        else:
            d.kind = 0
            d.strike, d.exercise = kw["strike"], kw["exerciseDate"]
            st = kw.get("settlementDate")
            d.settlement = st if st else kw["exerciseDate"]
        self.d = d

    def spec(self):
        kw, d = self.kw, self.d
        if d.kind == 1:
            return "product barrier %r %r %r %r %r %d" % (d.strike, d.barrier, d.maturity, d.monitor_freq, d.smooth,
                                                           d.call_put)
        if d.kind == 2:
            return "product europeans %d %s %s" % (d.n_options, " ".join(repr(float(x)) for x in self._keep[0]),
                                                    " ".join(repr(float(x)) for x in self._keep[1]))
        return "product european %r %r %r" % (d.strike, d.exercise, d.settlement)


def _check(rc):
    if rc < 0:
        raise RuntimeError(lib().or_last_error().decode())
    return rc


def num_payoffs(prd):
    return _check(lib().or_num_payoffs(ctypes.byref(prd.d)))


def num_params(mdl):
    return lib().or_num_params(ctypes.byref(mdl.d))


def timeline(mdl, prd):
    n = _check(lib().or_timeline(ctypes.byref(mdl.d), ctypes.byref(prd.d), None, None, 0))
    t = np.zeros(n)
    c = np.zeros(n, dtype=np.int32)
    lib().or_timeline(ctypes.byref(mdl.d), ctypes.byref(prd.d), t.ctypes.data_as(_dp),
                      c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n)
    return t, c.astype(bool)


def sobol_dirs(dims):
    out = np.zeros(32 * dims, dtype=np.uint32)
    lib().or_sobol_dirs(dims, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    return out.reshape(32, dims)


def sobol_points(dims, first, count):
    """uint32 points of paths first .. first + count - 1 (path p uses Sobol index p + 1)."""
    out = np.zeros((count + 1) * dims, dtype=np.uint32)
    _check(lib().or_sobol_states(dims, first, count, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))))
    return out.reshape(count + 1, dims)[1:]


def inv_normal(u):
    u = _arr(u)
    out = np.zeros_like(u)
    lib().or_inv_normal(len(u), u.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def value_paths(mdl, prd, n, first=0, mrg=None):
    """Per-path payoffs [n][nPay] of paths first .. first + n - 1 (mcSimul);
    Sobol, or mrg32k3a with mrg = (seed1, seed2) (even first path)."""
    npay = num_payoffs(prd)
    out = np.zeros(n * npay)
    s1, s2 = mrg if mrg else (0, 0)
    _check(lib().or_value_rng(ctypes.byref(mdl.d), ctypes.byref(prd.d), first, n, int(mrg is not None), s1, s2,
                              out.ctypes.data_as(_dp)))
    return out.reshape(n, npay)


def aad(mdl, prd, n, risk_index=0, weights=None, mrg=None):
    """mcSimulAAD over n paths: (payoffs [n][nPay], aggregated [n], risks [nParam])."""
    npay, npar = num_payoffs(prd), num_params(mdl)
    pay = np.zeros(n * npay)
    agg = np.zeros(n)
    risks = np.zeros(npar)
    w = _arr(weights) if weights is not None else None
    if w is not None and w.size != npay:
        raise ValueError("weights must have one entry per payoff (%d), got %d" % (npay, w.size))
    s1, s2 = mrg if mrg else (0, 0)
    _check(lib().or_aad_rng(ctypes.byref(mdl.d), ctypes.byref(prd.d), n, risk_index,
                            w.ctypes.data_as(_dp) if w is not None else None, int(mrg is not None), s1, s2,
                            pay.ctypes.data_as(_dp), agg.ctypes.data_as(_dp), risks.ctypes.data_as(_dp)))
    return pay.reshape(n, npay), agg, risks


# ---------------------------------------------------------------- reference driver

def ref_available():
    return os.path.exists(REF_DRIVER)


def ref_run(args, stdin=""):
    """Run the reference driver; returns its JSON output."""
    res = subprocess.run([REF_DRIVER] + [str(a) for a in args], input=stdin, capture_output=True, text=True,
                         check=True)
    return json.loads(res.stdout)


def ref_value(mdl, prd, n, threads=1, repeats=1, n_out=0, mrg=None):
    spec = mdl.spec() + "\n" + prd.spec() + "\n"
    if mrg:
        spec += "rng mrg %d %d\n" % tuple(mrg)
    return ref_run(["value", n, threads, repeats, n_out], spec)


def ref_aad(mdl, prd, n, risk_index=-1, threads=1, repeats=1, notionals=None, mrg=None):
    spec = mdl.spec() + "\n" + prd.spec() + "\n"
    if mrg:
        spec += "rng mrg %d %d\n" % tuple(mrg)
    if notionals:
        spec += "notionals %d %s\n" % (len(notionals), " ".join("%d %r" % (i, w) for i, w in notionals.items()))
    return ref_run(["aad", n, risk_index, threads, repeats, 0], spec)
