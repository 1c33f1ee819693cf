"""Writes golden per-pencil vectors from the reference's own Reconstruct / State classes.

Run in a container that has the reference sources (make -C oracle ref first):
    python tests/golden/make_golden.py
Inputs are seeded synthetic pencils (tests/cases.py); expected outputs come from
oracle/_ref/libref_{euler,ppm}.so, i.e. from reconstruct.cpp and euler/state.cpp compiled where
they lie, driven the way Grid::compute_x_flux drives them (grid/grid.cpp:225-241).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import cases  # noqa: E402
import orc  # noqa: E402


def main():
    out = {}
    for kind in ("euler", "ppm"):
        L = orc.ref_lib(kind)
        if L is None:
            raise SystemExit("oracle/_ref/libref_%s.so missing: run `make -C oracle ref`" % kind)
        rng = np.random.default_rng(20261203 if kind == "euler" else 20261204)
        P = cases.random_pencils(rng, 100, 12)
        F = np.full((100, 3, 12, 6), np.nan)
        E = np.full((100, 3, 12, 2), np.nan)
        for p in range(100):
            for d in range(3):
                f = np.full((12, 6), np.nan)
                e = np.full(24, np.nan)
                L.ref_pencil_flux(d, orc.dptr(np.ascontiguousarray(P[p])), orc.dptr(f), orc.dptr(e))
                F[p, d] = f
                E[p, d] = e.reshape(12, 2)
        out[kind + "_q"] = P
        out[kind + "_F"] = F
        out[kind + "_eig"] = E
    # floor and outflow on single cells
    L = orc.ref_lib("euler")
    rng = np.random.default_rng(99)
    cells = cases.random_pencils(rng, 20, 12).reshape(-1, 6)
    thr = cells[:40].copy()
    thr[:, 4] = 0.5 * (thr[:, 1] ** 2 + thr[:, 2] ** 2 + thr[:, 3] ** 2) / thr[:, 0] * 1.1111
    cells = np.concatenate([cells, thr])
    fl = cells.copy()
    for u in fl:
        L.ref_floor(orc.dptr(u))
    of = np.repeat(cells[:60, None, :], 6, axis=1).copy()
    for c in range(60):
        for face in range(6):
            L.ref_outflow(face, orc.dptr(of[c, face]))
    out["cells"] = cells
    out["floor"] = fl
    out["outflow"] = of
    np.savez_compressed(os.path.join(HERE, "ref_pencils.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_pencils.npz"))


if __name__ == "__main__":
    main()
