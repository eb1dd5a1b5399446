#!/usr/bin/env python3
"""Writes tests/golden/params_v1.wmhp (a kind-1 parameter file with 4 x 5
tables from a fixed numpy recipe) and tests/golden/params_v1_seeded.wmhp (a
kind-2 file: dim 1000, samples 64, seed 7), with the host writer of the
parameter-file format.  The committed files pin format version 1: later
versions of the engine must keep reading them.

Run from the repository root:  python3 tests/golden/make_params_file.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from minhashcuda_amd import params_file  # noqa: E402


def recipe_tables():
    """rs ~ Gamma(2,1), ln_cs = ln Gamma(2,1), betas ~ U[0,1): 4 samples x 5 dims."""
    rng = np.random.RandomState(3)
    rs = rng.gamma(2.0, 1.0, (4, 5)).astype(np.float32)
    ln_cs = np.log(rng.gamma(2.0, 1.0, (4, 5))).astype(np.float32)
    betas = rng.uniform(0.0, 1.0, (4, 5)).astype(np.float32)
    return rs, ln_cs, betas


if __name__ == "__main__":
    params_file.write(os.path.join(HERE, "params_v1.wmhp"), *recipe_tables())
    params_file.write_seeded(os.path.join(HERE, "params_v1_seeded.wmhp"), 1000, 64, 7)
    print("wrote params_v1.wmhp, params_v1_seeded.wmhp")
