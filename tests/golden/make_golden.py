"""Writes tests/golden/*.json: known-answer vectors computed with exact rational
arithmetic (fractions.Fraction), independent of the oracle's code.

The reference (BAOrec.jl, Julia) ships no unit tests or golden vectors for
its kernels (its test/runtests.jl testset is empty), so these vectors are
derived from the reference's formulas on inputs chosen so that every
intermediate is exactly representable in FP32:

* cic.json      CIC weights of single particles (src/mas.jl:1-52): cell
                indices and the 8 weights, including the x0 == n+1 and the
                wrap (x1 = 1) special cases.
* read_cic.json trilinear read-back values (src/mas.jl:271-318) of a small
                integer-valued field at dyadic positions.
* restrict.json 27-point full weighting (src/multigrid.jl:520-584) and
                trilinear prolongation (src/multigrid.jl:398-453) of small
                integer arrays, evaluated exactly.
* catalog.txt / catalog.json
                an ASCII catalog in the drivers' format (5 whitespace-separated
                columns, test/benchmarks_cpu.jl:22-23) with awkward decimal
                spellings, and the Float32 each value must parse to: the
                nearest float32 of the exact decimal (ties to even), found by
                rational arithmetic, not by a float parser.

Run:  python tests/golden/make_golden.py
"""
import json
import os
from fractions import Fraction as Fr

HERE = os.path.dirname(os.path.abspath(__file__))


def cic_case(p, n, L, b, w, wrap):
    """Exact CIC deposit of one particle: {(i,j,k) 0-based: weight}."""
    out_i0, out_i1, w0, w1 = [], [], [], []
    for a in range(3):
        u = (Fr(p[a]) - Fr(b[a])) * n[a] / Fr(L[a]) + 1
        x0 = u.numerator // u.denominator  # floor
        f1 = u - x0
        f0 = 1 - f1
        if x0 == n[a] + 1:
            x0 = 1
        x1 = 1 if (x0 == n[a] and wrap) else x0 + 1
        out_i0.append(x0 - 1)
        out_i1.append(x1 - 1)
        w0.append(f0)
        w1.append(f1)
    w0[0] *= Fr(w)
    w1[0] *= Fr(w)
    cells = {}
    for cx, wx in ((out_i0[0], w0[0]), (out_i1[0], w1[0])):
        for cy, wy in ((out_i0[1], w0[1]), (out_i1[1], w1[1])):
            for cz, wz in ((out_i0[2], w0[2]), (out_i1[2], w1[2])):
                key = (cx, cy, cz)
                cells[key] = cells.get(key, 0) + wx * wy * wz
    return cells


def gather_exact(field, p, n, L, b):
    d, u, wd, wu = [], [], [], []
    for a in range(3):
        dist = (Fr(p[a]) - Fr(b[a])) * n[a] / Fr(L[a])
        i = dist.numerator // dist.denominator
        fu = dist - i
        id_ = i + 1
        if id_ > n[a]:
            id_ -= n[a]
        iu = id_ + 1
        if iu > n[a]:
            iu -= n[a]
        d.append(id_ - 1)
        u.append(iu - 1)
        wd.append(1 - fu)
        wu.append(fu)
    tot = Fr(0)
    for ix, wx in ((d[0], wd[0]), (u[0], wu[0])):
        for iy, wy in ((d[1], wd[1]), (u[1], wu[1])):
            for iz, wz in ((d[2], wd[2]), (u[2], wu[2])):
                tot += Fr(field[iz][iy][ix]) * wx * wy * wz
    return tot


def restrict_exact(v, nf):
    nc = [x // 2 for x in nf]
    out = [[[Fr(0)] * nc[0] for _ in range(nc[1])] for _ in range(nc[2])]
    for kz in range(nc[2]):
        for ky in range(nc[1]):
            for kx in range(nc[0]):
                c = (2 * kx + 1, 2 * ky + 1, 2 * kz + 1)
                s = Fr(0)
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            m = abs(dx) + abs(dy) + abs(dz)
                            wgt = {0: 8, 1: 4, 2: 2, 3: 1}[m]
                            x = (c[0] + dx) % nf[0]
                            y = (c[1] + dy) % nf[1]
                            z = (c[2] + dz) % nf[2]
                            s += wgt * Fr(v[z][y][x])
                out[kz][ky][kx] = s / 64
    return out


def prolong_exact(v, nc):
    nf = [2 * x for x in nc]
    out = [[[None] * nf[0] for _ in range(nf[1])] for _ in range(nf[2])]
    for iz in range(nf[2]):
        for iy in range(nf[1]):
            for ix in range(nf[0]):
                idx = []
                for a, i in enumerate((ix, iy, iz)):
                    if i % 2 == 1:
                        idx.append([(i - 1) // 2])
                    else:
                        c = (i // 2 - 1) % nc[a]
                        idx.append([c, (c + 1) % nc[a]])
                vals = [Fr(v[z][y][x]) for z in idx[2] for y in idx[1] for x in idx[0]]
                out[iz][iy][ix] = sum(vals) / len(vals)
    return out


def nearest_f32(text):
    """Bits of the float32 nearest to the decimal `text` (ties to even)."""
    import struct
    import numpy as np
    t = text.lstrip("+")
    q = Fr(t)
    c = np.float32(float(q))
    best = None
    with np.errstate(over="ignore"):
        cands = (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf)))
    for cand in cands:
        if not np.isfinite(cand):
            continue
        d = abs(Fr(float(cand)) - q)
        bits = struct.unpack("<I", struct.pack("<f", float(cand)))[0]
        key = (d, bits & 1)
        if best is None or key < best[0]:
            best = (key, bits)
    if q == 0 and t.startswith("-"):
        return 0x80000000          # negative zero keeps its sign
    return best[1]


CATALOG_LINES = [
    "# ra dec z w nz  (comment lines and blank lines are skipped)",
    "10.5 -3.25 0.85 1 0.0004",
    "359.99997 -89.999997 0.9999999 1.0000001 3.0517578125e-05",
    "",
    "+0.1 0.2 0.3 .5 5.",
    "1e-3 2E2 7.006492321624085e-46 1.17549435e-38 3.4028234e38",
    "123.456789012345678 -0.0 0.099999994 0.100000001 0.10000000149011612",
    "   42\t43   44\t 45 46   ",
    "0.0000001 16777217 16777219 8388607.5 2.5e-45",
]


def main():
    cases = []
    n, L, b = (8, 8, 8), (8.0, 8.0, 8.0), (0.0, 0.0, 0.0)
    specs = [
        ((2.25, 3.5, 4.75), 1.0, True),      # generic interior
        ((0.0, 0.0, 0.0), 2.0, True),        # exactly on a node
        ((7.5, 7.25, 0.125), 1.0, True),     # last cell: x1 wraps to cell 0
        ((8.0, 1.5, 2.5), 1.0, True),        # p - b == L: x0 == n+1 -> cell 0
        ((3.0, 5.0, 6.5), 0.5, False),       # no wrap, interior
    ]
    for p, w, wrap in specs:
        cells = cic_case(p, n, L, b, w, wrap)
        cases.append({"p": p, "w": w, "wrap": wrap, "n": n, "L": L, "b": b,
                      "cells": [[k[0], k[1], k[2], float(v)] for k, v in sorted(cells.items()) if v != 0]})
    # a box with non-zero min and non-power-of-two grid
    n2, L2, b2 = (10, 6, 4), (20.0, 12.0, 8.0), (-4.0, 2.0, 1.0)
    for p in ((-3.5, 2.5, 1.5), (15.0, 13.0, 8.75)):
        cells = cic_case(p, n2, L2, b2, 1.0, True)
        cases.append({"p": p, "w": 1.0, "wrap": True, "n": n2, "L": L2, "b": b2,
                      "cells": [[k[0], k[1], k[2], float(v)] for k, v in sorted(cells.items()) if v != 0]})
    with open(os.path.join(HERE, "cic.json"), "w") as fh:
        json.dump(cases, fh, indent=1)

    field = [[[(3 * x + 5 * y + 7 * z + x * y * z) % 11 for x in range(8)] for y in range(8)] for z in range(8)]
    pts = [(0.5, 0.25, 0.125), (7.75, 7.5, 7.25), (3.0, 4.0, 5.0), (6.125, 0.875, 2.5)]
    gat = [{"p": p, "value": float(gather_exact(field, p, (8, 8, 8), (8.0, 8.0, 8.0), (0.0, 0.0, 0.0)))}
           for p in pts]
    with open(os.path.join(HERE, "read_cic.json"), "w") as fh:
        json.dump({"field": field, "n": [8, 8, 8], "L": [8.0] * 3, "b": [0.0] * 3, "points": gat}, fh)

    nf = (8, 6, 4)
    fine = [[[(x * 7 + y * 3 + z * 5) % 13 - 6 for x in range(nf[0])] for y in range(nf[1])] for z in range(nf[2])]
    coarse_in = [[[(x * 5 + y * 2 + z * 3) % 7 - 3 for x in range(4)] for y in range(3)] for z in range(2)]
    with open(os.path.join(HERE, "restrict.json"), "w") as fh:
        json.dump({
            "fine": fine, "fine_n": nf,
            "restricted": [[[float(v) for v in r] for r in pl] for pl in restrict_exact(fine, nf)],
            "coarse": coarse_in, "coarse_n": [4, 3, 2],
            "prolonged": [[[float(v) for v in r] for r in pl] for pl in prolong_exact(coarse_in, (4, 3, 2))],
        }, fh)


def write_catalog():
    lines = [ln.replace("\\t", "\t") for ln in CATALOG_LINES]
    with open(os.path.join(HERE, "catalog.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    rows = [[nearest_f32(tok) for tok in ln.split()] for ln in lines if ln.strip() and not ln.lstrip().startswith("#")]
    with open(os.path.join(HERE, "catalog.json"), "w") as fh:
        json.dump({"rows_f32_bits": rows}, fh)


if __name__ == "__main__":
    main()
    write_catalog()
