"""Seeded job mixes of the batching and serving workloads.

batch_jobs(n):   random CSG trees (depth 1-4, tape.random_csg) with resolutions drawn from
                 {32, 64, 96, 128, 192, 256} with weights {4, 4, 3, 3, 2, 1}; bounds 2;
                 reference parameters (96 -> 128 and 192 -> 256 by the reference's
                 power-of-two rounding).
serve_jobs(n):   the request mix of the serving workload: resolution log-uniform in
                 [32, 512] (rounded up to a power of two as the reference does), 80 %
                 random CSG trees and 20 % Mandelbulbs with 4-8 iterations (bounds 2.5).
poisson_arrivals(n, rate): arrival times of a Poisson process.
"""
from __future__ import annotations

import math

from . import Params
from . import tape

BATCH_RES = (32, 64, 96, 128, 192, 256)
BATCH_WEIGHTS = (4, 4, 3, 3, 2, 1)


def batch_jobs(n: int = 256, seed: int = 1):
    rng = tape.Rng(0xBA7C4 + seed)
    nodes, params = [], []
    for i in range(n):
        res = rng.choice_weighted(BATCH_RES, BATCH_WEIGHTS)
        nodes.append(tape.random_csg(seed * 100003 + i + 1))
        params.append(Params.reference(res, 2.0))
    return nodes, params


def serve_jobs(n: int, seed: int = 1):
    rng = tape.Rng(0x5E7E + seed)
    nodes, params = [], []
    for i in range(n):
        res = int(round(math.exp(rng.uniform(math.log(32), math.log(512)))))
        if rng.uniform(0.0, 1.0) < 0.8:
            nodes.append(tape.random_csg(seed * 7919 + i + 1))
            params.append(Params.reference(res, 2.0))
        else:
            nodes.append(tape.mandelbulb(4 + rng.below(5)))
            params.append(Params.reference(res, 2.5))
    return nodes, params


def poisson_arrivals(n: int, rate: float, seed: int = 1):
    rng = tape.Rng(0xA77 + seed)
    t, out = 0.0, []
    for _ in range(n):
        t += -math.log(1.0 - rng.uniform(0.0, 1.0)) / rate
        out.append(t)
    return out
