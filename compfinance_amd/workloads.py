"""Synthetic workloads of the benchmark configurations.

C1  Black-Scholes European call, S0=100, sigma=0.2, r=q=0, K=100, T=1, 2^16 paths
C2  Dupire local vol, up-and-out call K=120, B=150, 5% smoothing, 3y weekly
    (156 steps), 2^19 paths
C3  as C2 with price, delta and every local-vol sensitivity by AAD, 2^20 paths
C4  book on shared paths, Dupire 30x36 grid, 3y weekly (156 steps), 2^20 paths:
    30 European calls K = 70, 72.5, .., 142.5; 10 up-and-out calls K evenly
    100..130 with barriers 140, 145, .., 185 (5% smoothing, weekly); 5 Asian
    calls K = 90, 95, .., 110 on monthly fixings moved to the nearest weekly
    date (keeps 156 steps); notionals uniform [0.5, 2] (numpy PCG64, seed
    20261203); aggregated risk to 1,081 parameters
C5  serving mix (see serving.py)

Local-vol grid: spots_i = 50 + 150 i / (nS - 1), times_j = (j + 1) / 12,
j = 0..35, sigma(S, t) = 0.15 + 0.10 ln(S/100)^2 + 0.05 exp(-t).  The headline
grid has nS = 30 (1,080 vols, the reference's own demo), the literal one
nS = 60 (2,160 vols).
"""
import math

import numpy as np

WEEK = 1.0 / 52.0


def lv_grid(n_spots=30, n_times=36):
    spots = np.array([50.0 + 150.0 * i / (n_spots - 1) for i in range(n_spots)])
    times = np.array([(j + 1) / 12.0 for j in range(n_times)])
    vols = np.empty((n_spots, n_times))
    for i, s in enumerate(spots):
        for j, t in enumerate(times):
            vols[i, j] = 0.15 + 0.10 * math.log(s / 100.0) ** 2 + 0.05 * math.exp(-t)
    return spots, times, vols


def c1():
    """(model args, product args) of C1 for BlackScholes / European."""
    return dict(spot=100.0, vol=0.20, spotMeasure=False, rate=0.0, div=0.0), dict(strike=100.0, exerciseDate=1.0)


def dupire_args(n_spots=30, n_times=36, spot=100.0, max_dt=WEEK):
    spots, times, vols = lv_grid(n_spots, n_times)
    return dict(spot=spot, spots=spots, times=times, vols=vols, maxDt=max_dt)


def barrier_args(strike=120.0, barrier=150.0, maturity=3.0, freq=WEEK, smooth=0.05, put=False):
    return dict(strike=strike, barrier=barrier, maturity=maturity, monitorFreq=freq, smooth=smooth, callPut=put)


def c4_parts():
    """C4 book parts as (kind, keyword arguments), kind in european(s),
    barrier, asian; the keywords fit both the engine classes and the oracle."""
    parts = [("europeans", dict(maturities=[3.0] * 30, strikes=[70.0 + 2.5 * i for i in range(30)]))]
    for i in range(10):
        parts.append(("barrier", barrier_args(strike=100.0 + 30.0 * i / 9, barrier=140.0 + 5.0 * i)))
    for k in (90.0, 95.0, 100.0, 105.0, 110.0):
        parts.append(("asian", dict(strike=k, maturity=3.0, fixingFreq=1.0 / 12, snapFreq=WEEK, callPut=False)))
    return parts


def c4_notionals(n_pay, seed=20261203):
    return np.random.default_rng(seed).uniform(0.5, 2.0, n_pay)


C1_PATHS = 1 << 16
C2_PATHS = 1 << 19
C3_PATHS = 1 << 20
C4_PATHS = 1 << 20

# Black-Scholes closed form of C1 (analytics.h:8-20 with r = q = 0)
C1_CLOSED_FORM = 7.965567455405804
