#!/usr/bin/env python3
"""Writes the launch-shape table (tcx_autotune, see include/timesc.h) for one
MI355X: the n=100000, m=512, k=16 fit and, for every FFT length of series up
to 4096 and k up to 16 and up to 32, a fit of 2500 series of length 3F/8
(the middle of the lengths that take that F) on the benchmark generator.
Each entry only selects among launch shapes that give identical results.

    python times-c_amd/tune_mi355x.py [OUT]   (default: tune_mi355x.txt here)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import timesc as T  # noqa: E402


def shapes():
    yield 100000, 512, 16
    for F in (128, 256, 512, 1024, 2048, 4096, 8192):
        m = 3 * F // 8
        for k in (8, 24):
            yield 2500, m, k


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "tune_mi355x.txt")
    T.tune_reset()
    ctx = T.Context(0)
    try:
        for n, m, k in shapes():
            X, _ = T.generate(11, n, m, k, 0.3)
            d = ctx.to_device(X)
            try:
                ctx.znorm(d, ddof=0)
                rep = ctx.autotune(d, k, reps=3)
            finally:
                d.free()
            print(f"n={n} m={m} k={k}: {rep}", flush=True)
    finally:
        ctx.close()
    T.tune_save(out)
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
