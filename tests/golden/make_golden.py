"""Writes tests/golden/*.npz from the reference's own compiled C stage 1.

The reference ships no golden vectors (it has no tests).  Its single-thread C stage 1,
src/gravitation/kernel/_lib1_/lib.c, is compiled from the reference tree into oracle/_ref/ by
oracle/Makefile; this script runs it on seeded inputs and stores inputs and outputs, so that
the fixtures travel to hosts without the reference tree.

    make -C oracle && python tests/golden/make_golden.py

Fixtures
--------
c1_lib1.npz: the 1024-body uniform-cube case (seed 0, side 1e3, masses U[1e-3, 1e-1],
  G = 6.6743e-11): pos (1024, 3) float32, m (1024,) float32, G, and acc (1024, 3) float32 from
  lib1's step_stage1 (FP32, -O3 -ffast-math, flattened unique-pair loop, lib.c:52-126).
solarsystem_lib1.npz: the reference's two-body scenario (src/gravitation/lib/simulation.py:
  85-97) with the base G (src/gravitation/kernel/_base_.py:73): pos, m, G, acc from lib1.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref as R  # noqa: E402


def main():
    if not R.available('lib1'):
        raise SystemExit('oracle/_ref/lib1.so missing: run `make -C oracle` with the reference '
                         'tree present')
    rng = np.random.default_rng(0)
    pos = rng.uniform(0.0, 1e3, (1024, 3)).astype(np.float32)
    m = rng.uniform(1e-3, 1e-1, 1024).astype(np.float32)
    G = 6.6743e-11
    acc = R.RefLib1(pos, m, G).accel()
    np.savez(os.path.join(HERE, 'c1_lib1.npz'), pos=pos, m=m, G=np.float64(G), acc=acc)

    au = 149597870700.0
    pos2 = np.array([[0.0, 0.0, 0.0], [0.0, -au, 0.0]], dtype=np.float32)
    m2 = np.array([1.98892e30, 5.97237e24], dtype=np.float32)
    G2 = 6.6740831e-11
    acc2 = R.RefLib1(pos2, m2, G2).accel()
    np.savez(os.path.join(HERE, 'solarsystem_lib1.npz'), pos=pos2, m=m2, G=np.float64(G2),
             acc=acc2)
    print('wrote c1_lib1.npz, solarsystem_lib1.npz')


if __name__ == '__main__':
    main()
