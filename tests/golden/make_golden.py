"""Writes the golden fixtures for the Green's tensors.

The point tensor is derived here independently of the closed form the
reference and the oracle use (greens.rs:34-71): it is computed as
(k^2 I + grad grad) e^{ikR} / (4 pi R) by mpmath numerical differentiation
of the scalar Green's function at 40 significant digits. A wrong sign, a
dropped term or a transposed index in the closed form cannot match it.

The FCD tensor (greens.rs:94-167) is the Gauss-Legendre 3-point average of
that same derived tensor over the source cell, with nodes and weights
computed from the Legendre polynomial P3 (roots of x(5x^2-3)/2, weights
2/((1-x^2) P3'(x)^2)), not typed in.

Run: python tests/golden/make_golden.py  (takes ~1 minute)
"""
import json
import os

import mpmath as mp

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def scalar_g(k, x, y, z):
    R = mp.sqrt(x * x + y * y + z * z)
    return mp.exp(1j * k * R) / (4 * mp.pi * R)


def dyadic_by_derivatives(k, d):
    k = mp.mpf(k)
    x0, y0, z0 = (mp.mpf(v) for v in d)
    G = [[None] * 3 for _ in range(3)]
    for a in range(3):
        for b in range(3):
            order = [0, 0, 0]
            order[a] += 1
            order[b] += 1
            dd = mp.diff(lambda x, y, z: scalar_g(k, x, y, z), (x0, y0, z0), tuple(order))
            G[a][b] = dd + (k * k * scalar_g(k, x0, y0, z0) if a == b else 0)
    return G


def gl3():
    # P3(x) = (5x^3 - 3x)/2, P3'(x) = (15x^2 - 3)/2
    roots = sorted(mp.polyroots([mp.mpf(5) / 2, 0, mp.mpf(-3) / 2, 0]), key=lambda r: mp.re(r))
    nodes = [mp.re(r) for r in roots]
    weights = [2 / ((1 - x * x) * ((15 * x * x - 3) / 2) ** 2) for x in nodes]
    return nodes, weights


def fcd_by_quadrature(k, ro, rs, cell):
    nodes, weights = gl3()
    half = mp.mpf(cell) / 2
    acc = [[mp.mpc(0)] * 3 for _ in range(3)]
    tw = mp.mpf(0)
    for ix in range(3):
        for iy in range(3):
            for iz in range(3):
                rp = (mp.mpf(rs[0]) + half * nodes[ix], mp.mpf(rs[1]) + half * nodes[iy],
                      mp.mpf(rs[2]) + half * nodes[iz])
                d = [mp.mpf(ro[t]) - rp[t] for t in range(3)]
                if d[0] ** 2 + d[1] ** 2 + d[2] ** 2 < mp.mpf("1e-20"):
                    continue
                w = weights[ix] * weights[iy] * weights[iz]
                G = dyadic_by_derivatives(k, d)
                for a in range(3):
                    for b in range(3):
                        acc[a][b] += w * G[a][b]
                tw += w
    return [[acc[a][b] / tw for b in range(3)] for a in range(3)]


def enc(G):
    return [[[float(mp.re(G[a][b])), float(mp.im(G[a][b]))] for b in range(3)] for a in range(3)]


def main():
    ks = [2 * mp.pi * 1.33 / 520, 2 * mp.pi / 400, 2 * mp.pi * 1.33 / 800]
    disps = [(10.0, 5.0, 3.0), (2.0, 0.0, 0.0), (0.0, 2.0, 0.0), (1.5, 1.5, 1.5),
             (-3.0, 4.5, 0.0), (40.0, -20.0, 12.0), (0.7, -0.2, 0.3), (95.0, 38.0, 38.0),
             (-6.0, 0.0, -6.0), (300.0, 0.0, 0.0), (12.5, -33.25, 7.75)]
    point = []
    for k in ks:
        for d in disps:
            point.append({"k": float(k), "r1": list(d), "r2": [0.0, 0.0, 0.0],
                          "G": enc(dyadic_by_derivatives(k, d))})
    fcd = []
    cases = [((0.0, 0.0, 0.0), (2.0, 0.0, 0.0), 2.0), ((0.0, 0.0, 0.0), (2.0, 2.0, 0.0), 2.0),
             ((0.0, 0.0, 0.0), (4.0, 0.0, 0.0), 2.0), ((0.0, 0.0, 0.0), (1.5, 1.5, 1.5), 1.5),
             ((1.0, 2.0, 3.0), (1.0, 5.0, 3.0), 3.0), ((0.0, 0.0, 0.0), (3.0, 0.0, 0.0), 1.5)]
    for k in ks[:2]:
        for ro, rs, cell in cases:
            fcd.append({"k": float(k), "r_obs": list(ro), "r_src": list(rs), "cell": cell,
                        "G": enc(fcd_by_quadrature(k, ro, rs, cell))})
    with open(os.path.join(HERE, "greens_point.json"), "w") as f:
        json.dump(point, f, indent=0)
    with open(os.path.join(HERE, "greens_fcd.json"), "w") as f:
        json.dump(fcd, f, indent=0)
    print(len(point), "point tensors,", len(fcd), "FCD tensors")


if __name__ == "__main__":
    main()
