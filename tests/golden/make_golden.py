"""Writes the golden fixtures under tests/golden/ from textbook library
routines (SciPy erfcinv/erfc, NumPy leggauss); nothing here reads or runs the
reference.  Re-run with:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np
from scipy import special

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ys = [1e-12, 1e-9, 1e-6, 1e-5, 9e-5, 1e-4, 4.5e-4, 1e-3, 0.01, 0.1, 0.5,
          1.0, 1.5, 1.9]
    erfc = {"y": ys, "x": [float(special.erfcinv(y)) for y in ys]}
    gl = {}
    for n in (2, 4, 6, 12, 24, 38, 52, 68, 100, 160, 250, 324, 3, 7, 21, 41):
        x, w = np.polynomial.legendre.leggauss(n)
        keep = x >= -1e-300 if n % 2 else x > 0
        # non-negative roots ascending (0 first for odd n)
        xs, ws = x[keep], w[keep]
        if n % 2:
            xs[0] = 0.0
        gl[str(n)] = {"x": xs.tolist(), "w": ws.tolist()}
    with open(os.path.join(HERE, "erfc_inv.json"), "w") as f:
        json.dump(erfc, f, indent=1)
    with open(os.path.join(HERE, "gauss_legendre.json"), "w") as f:
        json.dump(gl, f)


if __name__ == "__main__":
    main()
