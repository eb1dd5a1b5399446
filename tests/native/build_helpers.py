"""Builds the test helpers in this folder (test infrastructure, not part of
the engine): libfastdiv_check.so from fastdiv_check.hip, compiled with the
engine's own device flags so that the division it checks is the engine's, and
libstripe_check.so from stripe_check.cpp, the persistent kernel's stripe-piece
search compiled as host code with the engine's host flags.

Usage:  python tests/native/build_helpers.py   (or build() from Python)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libfastdiv_check.so")
STRIPE_LIB = os.path.join(HERE, "libstripe_check.so")


def build_stripe_check(verbose=False):
    sys.path.insert(0, ROOT)
    from minhashcuda_amd import build as engine
    cmd = ([engine.hipcc(), "-shared"] + engine.COMMON + engine.HOST +
           [os.path.join(HERE, "stripe_check.cpp"), "-o", STRIPE_LIB + ".tmp"])
    engine._run(cmd, verbose)
    os.replace(STRIPE_LIB + ".tmp", STRIPE_LIB)
    return STRIPE_LIB


def build(verbose=False):
    sys.path.insert(0, ROOT)
    from minhashcuda_amd import build as engine
    cmd = ([engine.hipcc(), "-shared"] + engine.COMMON + engine.DEVICE +
           [os.path.join(HERE, "fastdiv_check.hip"), "-o", LIB + ".tmp"])
    engine._run(cmd, verbose)
    os.replace(LIB + ".tmp", LIB)
    build_stripe_check(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
