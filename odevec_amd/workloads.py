"""Seeded synthetic workloads (NumPy PCG64) for the parity tests and the bench.

Each generator returns a dict with the network id, the state ``y`` (shape
(neq, ncells), cells fastest), per-cell ``t0``, ``t_stop``, ``dt0``, optional
``params`` (shape (nparam, ncells)) and the tolerances.

* ROBER (seed 1): k_j = k_j0 * exp(0.1 z), z ~ N(0,1) per cell and rate,
  k0 = (0.04, 3e7, 1e4); y0 = (1, 0, 0); t 0 -> 1e5; dt hint 1e-10
  (tests/rober.f90:58); rtol 1e-6, atol 1e-10.
* OREGO (seed 2): y0 = (1, 2, 3) * U[0.5, 1.5]; t 0 -> 360; dt hint 1e-10
  (tests/orego.f90:129); rtol 1e-6, atol 1e-10.
* PRIMORDIAL (seed 3): n_H ~ logU[1, 1e12] cm^-3, T ~ logU[10, 1e4] K,
  x(H2) = n(H2)/n_H ~ logU[1e-6, 1e-2], x(e) = n(e)/n_H ~ logU[1e-8, 1e-3];
  n(H+) = n(e) (charge balance), He/H = 0.0806 by number (the mass fractions
  0.24375/0.75615 of tests/primordial.f90:50,55 with KROME's masses),
  D/H = 2.5e-5, neutral H closes the hydrogen nuclei, trace ions at 1e-40
  (tests/primordial.f90:53-61), CR = g = dummy = 0, Tgas slot = T;
  t_stop ~ logU[1e8, 1e14] s; dt hint 2.2954626201190183e-26 as the single
  precision literal of tests/primordial.f90:73; rtol 1e-6, atol 1e-20,
  MinimalTimestep 1e-20 (tests/primordial.f90:44-45).
* SYNTH64 (serving): abundances logU[1e-6, 1] of n, n ~ logU[1, 1e6],
  T ~ logU[100, 1e4] K.
* Batching mix (seed 4): categorical 40/30/30 ROBER/OREGO/PRIMORDIAL cells,
  t_stop log-uniform over 6 decades per network (ROBER [1e-1, 1e5], OREGO
  [3.6e-4, 360], PRIMORDIAL [1e8, 1e14]), each network's own inputs and
  tolerances as above.
"""
from __future__ import annotations

import numpy as np

HE_PER_H = (0.24375 / 6.69206503638e-24) / (0.75615 / 1.67353251819e-24)
D_PER_H = 2.5e-5
FLOOR = 1e-40
PRIMORDIAL_DT_HINT = float(np.float32(2.2954626201190183e-26))


def _logu(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def rober(n: int, seed: int = 1, t_stop: float = 1e5, sigma: float = 0.1):
    rng = np.random.default_rng(seed)
    k0 = np.array([0.04, 3e7, 1e4])
    params = k0[:, None] * np.exp(sigma * rng.standard_normal((3, n)))
    y = np.zeros((3, n))
    y[0] = 1.0
    return dict(network="ROBER/NONE/DENSE", y=y, t0=np.zeros(n), t_stop=np.full(n, float(t_stop)),
                dt0=np.full(n, 1e-10), params=params, rtol=1e-6, atol=1e-10, dt_min=float(np.float32(1e-10)))


def orego(n: int, seed: int = 2, t_stop: float = 360.0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.5, 1.5, n)
    y = np.array([1.0, 2.0, 3.0])[:, None] * u[None, :]
    return dict(network="OREGO/NONE/DENSE", y=np.ascontiguousarray(y), t0=np.zeros(n),
                t_stop=np.full(n, float(t_stop)), dt0=np.full(n, 1e-10), params=None, rtol=1e-6, atol=1e-10,
                dt_min=float(np.float32(1e-10)))


def primordial_state(nH, T, xH2, xe):
    n = nH.size
    y = np.zeros((16, n))
    ne = xe * nH
    y[0] = ne                      # E
    y[1] = FLOOR                   # H-
    y[2] = nH * (1.0 - 2.0 * xH2 - xe)  # H
    y[3] = HE_PER_H * nH           # HE
    y[4] = xH2 * nH                # H2
    y[5] = D_PER_H * nH            # D
    y[6] = FLOOR                   # HD
    y[7] = ne                      # H+
    y[8] = FLOOR                   # HE+
    y[9] = FLOOR                   # H2+
    y[10] = FLOOR                  # D+
    y[11] = FLOOR                  # HE++
    y[14] = T                      # Tgas slot
    return y


def primordial(n: int, seed: int = 3, network: str = "PRIMORDIAL/CMK/CSC", t_lo: float = 1e8,
               t_hi: float = 1e14, rtol: float = 1e-6):
    rng = np.random.default_rng(seed)
    nH = _logu(rng, 1.0, 1e12, n)
    T = _logu(rng, 10.0, 1e4, n)
    xH2 = _logu(rng, 1e-6, 1e-2, n)
    xe = _logu(rng, 1e-8, 1e-3, n)
    ts = _logu(rng, t_lo, t_hi, n)
    return dict(network=network, y=primordial_state(nH, T, xH2, xe), t0=np.zeros(n), t_stop=ts,
                dt0=np.full(n, PRIMORDIAL_DT_HINT), params=None, rtol=rtol, atol=1e-20, dt_min=1e-20)


def synth64(n: int, seed: int = 7, network: str = "SYNTH64/FF/CSC", t_stop: float = 1e6, rtol: float = 1e-6):
    rng = np.random.default_rng(seed)
    ntot = _logu(rng, 1.0, 1e6, n)
    y = ntot[None, :] * np.exp(rng.uniform(np.log(1e-6), 0.0, (64, n)))
    T = _logu(rng, 100.0, 1e4, n)
    return dict(network=network, y=np.ascontiguousarray(y), t0=np.zeros(n), t_stop=np.full(n, float(t_stop)),
                dt0=np.full(n, 1e-10), params=T[None, :].copy(), rtol=rtol, atol=1e-20, dt_min=1e-20)


def subset(w: dict, idx) -> dict:
    out = dict(w)
    out["y"] = np.ascontiguousarray(w["y"][:, idx])
    for k in ("t0", "t_stop", "dt0"):
        out[k] = np.ascontiguousarray(w[k][idx])
    if w.get("params") is not None:
        out["params"] = np.ascontiguousarray(w["params"][:, idx])
    return out


BATCH_SPANS = {"ROBER/NONE/DENSE": (1e-1, 1e5), "OREGO/NONE/DENSE": (3.6e-4, 360.0),
               "PRIMORDIAL/CMK/CSC": (1e8, 1e14)}


def batch_mix(n: int, seed: int = 4, shares=(0.4, 0.3, 0.3)):
    """The batching workload: one segment per network (dicts as above), cell
    counts drawn categorically with `shares`, t_stop log-uniform over each
    network's 6-decade span."""
    rng = np.random.default_rng(seed)
    nets = list(BATCH_SPANS)
    counts = np.bincount(rng.choice(len(nets), size=n, p=np.asarray(shares) / np.sum(shares)), minlength=len(nets))
    gens = {"ROBER/NONE/DENSE": lambda m, sd: rober(m, seed=sd), "OREGO/NONE/DENSE": lambda m, sd: orego(m, seed=sd),
            "PRIMORDIAL/CMK/CSC": lambda m, sd: primordial(m, seed=sd)}
    segs = []
    for k, (net, m) in enumerate(zip(nets, counts)):
        w = gens[net](int(m), 100 * seed + k)
        lo, hi = BATCH_SPANS[net]
        w["t_stop"] = _logu(rng, lo, hi, int(m))
        segs.append(w)
    return segs
