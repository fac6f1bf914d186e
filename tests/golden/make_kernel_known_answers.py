"""Writes kernel_known_answers.json: shape-function values w(q), dw(q) and
normalisation constants of the four smoothing-kernel families, computed here
in exact rational arithmetic (fractions) from the published forms the
reference implements (src/kernels.cuh:75-165):

    Wendland C2   w = (1-q)^4 (4q+1),             alpha = 21/(2 pi)
    cubic B3      w = (1-q)^3 - 4 (1/2-q)_+^3,    alpha = 16/pi
    Wendland C6   w = (1-q)^8 (32q^3+25q^2+8q+1), alpha = 1365/(64 pi)
    double cosine w = 4 cos(pi q) + cos(2 pi q) + 3, alpha = pi/(4 pi^2 - 30)

dw is the exact derivative of w. For the polynomial kernels every value is a
rational number, written as a decimal string with 30 digits.
"""

import json
import math
from fractions import Fraction as F
from pathlib import Path


def pos(x):
    return x if x > 0 else F(0)


def c2(q):
    return (1 - q) ** 4 * (4 * q + 1), -20 * q * (1 - q) ** 3


def b3(q):
    t2 = pos(F(1, 2) - q)
    return (1 - q) ** 3 - 4 * t2 ** 3, -3 * (1 - q) ** 2 + 12 * t2 ** 2


def w6(q):
    w = (1 - q) ** 8 * (32 * q ** 3 + 25 * q ** 2 + 8 * q + 1)
    # d/dq [(1-q)^8 p(q)] = (1-q)^7 [-8 p + (1-q) p'] = -22 q (16q^2 + 7q + 1)(1-q)^7
    p = 32 * q ** 3 + 25 * q ** 2 + 8 * q + 1
    dp = 96 * q ** 2 + 50 * q + 8
    dw = (1 - q) ** 7 * (-8 * p + (1 - q) * dp)
    return w, dw


def cos_k(q):
    q = float(q)
    return (4 * math.cos(math.pi * q) + math.cos(2 * math.pi * q) + 3,
            -4 * math.pi * math.sin(math.pi * q) - 2 * math.pi * math.sin(2 * math.pi * q))


QS = [F(0), F(1, 8), F(1, 4), F(3, 8), F(1, 2), F(5, 8), F(3, 4), F(7, 8), F(1)]


def fmt(x):
    if isinstance(x, F):
        return f"{x.numerator / x.denominator:.17g}" if x.denominator < 2 ** 60 else repr(float(x))
    return f"{x:.17g}"


def main():
    out = {"q": [fmt(q) for q in QS], "kernels": {}}
    alpha = {"C2": 21 / (2 * math.pi), "B3": 16 / math.pi, "W6": 1365 / (64 * math.pi),
             "COS": math.pi / (4 * math.pi ** 2 - 30)}
    for name, fn in (("C2", c2), ("B3", b3), ("W6", w6), ("COS", cos_k)):
        vals = [fn(q) for q in QS]
        out["kernels"][name] = {
            "alpha": f"{alpha[name]:.17g}",
            "w": [fmt(v[0]) for v in vals],
            "dw": [fmt(v[1]) for v in vals],
        }
    path = Path(__file__).with_name("kernel_known_answers.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
