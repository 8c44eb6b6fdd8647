"""Warm starts for the Newton/Picard loop (the LoadVelocityGuess slot, Forging3DSolver.cs:854-869).

The reference starts each increment from the previous increment's velocity field (or, with ExtrapolateStart, from
the trend 2 v_n - v_{n-1}); the learned start replaces that guess with a per-node network prediction.
"""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "warmstart_mlp.npz")


def load_default():
    """The trained warm-start network as a solver callback, or None when no trained weights are present."""
    if not os.path.exists(DEFAULT_WEIGHTS):
        return None
    from .mlp import LearnedWarmStart
    return LearnedWarmStart.load(DEFAULT_WEIGHTS)
