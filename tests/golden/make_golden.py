"""Writes tests/golden/golden.json: closed-form values that pin the oracle
independently of its C code.

  key_geometry   src/structs.jl:63-82 evaluated with exact rational
                 arithmetic (fractions.Fraction of the decimal inputs); every
                 case is checked to sit far from a floor() boundary so the
                 FP64 quotient floors to the same integer.
  lattice_counts neighbours of an interior point of an infinite square /
                 cubic lattice within R = k*dr, by integer enumeration
                 (i^2 + j^2 (+ k^2) <= (R/dr)^2, origin excluded).
  kernel_consts  the Wendland normalisations of src/kernels.jl in closed form.

Run: python tests/golden/make_golden.py
"""
import json
import math
import os
from fractions import Fraction

CASES = [
    # (h, bmin, bmax) as decimal strings
    ("0.048", ("-0.07", "-0.07", "0"), ("4.07", "3.07", "0")),     # 2D dam break, R = 2.4 dr
    ("0.096", ("-0.07", "-0.07", "0"), ("4.07", "3.07", "0")),     # 2D dam break, R = 4.8 dr
    ("0.06", ("-0.04375", "-0.04375", "-0.04375"), ("4.04375", "1.04375", "3.04375")),  # 3D, R = 4.8 dr, dr = 0.0125
    ("0.03792", ("-0.02765", "-0.02765", "-0.02765"), ("4.02765", "1.02765", "3.02765")),  # 3D, dr = 0.0079
    ("0.048", ("-10", "-10", "0"), ("10", "10", "0")),             # collision test domain
    ("0.25", ("0", "0", "0"), ("0.4", "1", "0")),                  # tiny grid, L1 = 2
    ("0.3", ("-1", "0", "-1"), ("1", "0.5", "1")),                 # 3D, L2 = 2
]


def floor_frac(q):
    return math.floor(q)


def key_geometry(h, bmin, bmax):
    hf = Fraction(h)
    phase, lim = [], []
    for lo, hi in zip(bmin, bmax):
        qlo, qhi = Fraction(lo) / hf, Fraction(hi) / hf
        # guard: the FP64 quotient must floor like the exact one
        flo, fhi = float(lo) / float(h), float(hi) / float(h)
        for exact, approx in ((qlo, flo), (qhi, fhi)):
            assert math.floor(approx) == floor_frac(exact)
            if exact.denominator != 1:
                assert abs(float(exact) - round(float(exact))) > 1e-9
        p = floor_frac(qlo)
        phase.append(p)
        lim.append(floor_frac(qhi) - p + 1)
    key_max = lim[0] * lim[1] * lim[2]
    diff = []
    if lim[2] == 1:
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                diff.append(di + lim[0] * dj)
    else:
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    diff.append(di + lim[0] * (dj + lim[1] * dk))
    return dict(h=float(h), bmin=[float(v) for v in bmin], bmax=[float(v) for v in bmax],
                key_phase=phase, key_lim=lim, key_max=key_max, key_diff=diff)


def lattice_count(dim, k):
    r2 = Fraction(k) ** 2
    n = int(math.ceil(float(k)))
    cnt = 0
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            for l in (range(-n, n + 1) if dim == 3 else (0,)):
                s = i * i + j * j + l * l
                if 0 < s and s <= r2:
                    cnt += 1
    return cnt


def main():
    out = dict(
        key_geometry=[key_geometry(*c) for c in CASES],
        lattice_counts=[dict(dim=d, R_over_dr=float(k), count=lattice_count(d, Fraction(k)))
                        for d in (2, 3) for k in ("2.4", "4.8")],
        kernel_consts=dict(w2=7.0 / math.pi, w3=21.0 / (2.0 * math.pi),
                           dw2=140.0 / math.pi, dw3=210.0 / math.pi),
    )
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
