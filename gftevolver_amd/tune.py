"""Launch-shape autotuner for the momentum-form step kernel.

Every tile shape (bz x by), block order and LDS staging choice produces
bitwise-identical fields (tests/test_gpu_momentum.py checks this), so tuning
only picks the fastest: each candidate runs a few steps on the live state and
is timed by the library's HIP events on its own stream (gft_last_step_ms).
The FP32 SM kernel is compiled for 4 waves per SIMD by default; the other
occupancy variants (unconstrained 0, 3) are separately compiled kernels whose
results may differ in the last bit; they, and staging modes 1 and 0, are
tried only with occupancy=True.  A candidate is (bz, by, order, min_waves, stage).
"""
from __future__ import annotations

TILES = ((256, 1), (128, 2), (64, 4), (32, 8), (16, 16))
ORDERS = (3, 2, 1)


def candidates(occupancy: bool = False):
    return [(bz, by, order, mw, st) for (bz, by) in TILES for order in ORDERS
            for mw in ((4, 3, 0) if occupancy else (4,)) for st in ((2, 1, 0) if occupancy else (2,))]


def tune(ev, candidates_=None, steps: int = 6, rounds: int = 2, fp32_sm: bool = True, occupancy: bool = False):
    """Times each candidate on evolver `ev` (advancing its state) and leaves the
    fastest installed.  Returns (best, {candidate: ms})."""
    cands = list(candidates_ or candidates(occupancy))
    if not fp32_sm:
        # The occupancy variants exist for the FP32 SM kernel only.
        cands = sorted({(bz, by, r, 4, st) for (bz, by, r, _, st) in cands})
    timing = {}
    for _ in range(rounds):
        for c in cands:
            ev.set_launch(*c)
            ev.evolve(1)
            ev.evolve(steps)
            ms = ev.last_step_ms()[0]
            timing[c] = min(ms, timing.get(c, float("inf")))
    best = min(timing, key=timing.get)
    ev.set_launch(*best)
    return best, timing
