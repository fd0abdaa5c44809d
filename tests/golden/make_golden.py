"""Writes the small golden fixtures under tests/golden/.

The reference (milk-white-way/AMRESSIF) ships no test vectors for its time
step, so these fixtures come from routes independent of the oracle's code:

* poisson_dct_n16.npz -- mean-zero solution of the 16 x 16 homogeneous-Neumann
  5-point Poisson problem via scipy.fft.dctn / idctn (type II, orthonormal),
  using the closed-form eigenvalues -(4/h^2) sin^2(pi k / 2n) per direction.
* quick_cases.json -- QUICK face values of Leonard (1979) evaluated in exact
  rational arithmetic: face value = 6/8 upwind + 3/8 downwind - 1/8
  far-upwind, flux = face velocity x face value.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
from fractions import Fraction

import numpy as np
import scipy.fft

HERE = os.path.dirname(os.path.abspath(__file__))


def poisson_fixture(n=16, seed=4242):
    h = 1.0 / n
    rng = np.random.Generator(np.random.PCG64(seed))
    b = rng.standard_normal((n, n))
    b -= b.mean()
    bh = scipy.fft.dctn(b, type=2, norm="ortho")
    k = np.arange(n)
    lam1 = -4.0 / (h * h) * np.sin(np.pi * k / (2 * n)) ** 2
    lam = lam1[:, None] + lam1[None, :]
    lam[0, 0] = 1.0
    xh = bh / lam
    xh[0, 0] = 0.0
    x = scipy.fft.idctn(xh, type=2, norm="ortho")
    np.savez(os.path.join(HERE, "poisson_dct_n16.npz"), b=b, x=x, h=np.array(h))


def quick_fixture():
    cases = []
    stencils = [(1, 2, 4, 8), (0, 0, 1, 0), (3, -1, 2, 5), (-2, 7, 7, -3)]
    vels = [Fraction(1), Fraction(-1), Fraction(3, 2), Fraction(-5, 4), Fraction(0)]
    for s in stencils:
        a2, a1, b0, b1 = (Fraction(x) for x in s)  # U(i-2), U(i-1), U(i), U(i+1)
        for w in vels:
            if w > 0:   # upwind cell i-1, downwind i, far-upwind i-2
                face = Fraction(6, 8) * a1 + Fraction(3, 8) * b0 - Fraction(1, 8) * a2
            elif w < 0:  # upwind cell i, downwind i-1, far-upwind i+1
                face = Fraction(6, 8) * b0 + Fraction(3, 8) * a1 - Fraction(1, 8) * b1
            else:
                face = Fraction(0)
            cases.append({"U": [int(x) for x in s], "ucont": float(w), "flux": float(w * face)})
    with open(os.path.join(HERE, "quick_cases.json"), "w") as f:
        json.dump(cases, f, indent=1)


if __name__ == "__main__":
    poisson_fixture()
    quick_fixture()
    print("wrote fixtures to", HERE)
