"""Writes the golden fixtures: synthetic circulation problems and their exact binary64 solutions.

The reference stores no streamfunction vectors, so these fixtures come from an independent direct
solve: the finest grid's operator (the stencil's binary32 weights promoted to binary64) assembled
as a scipy sparse matrix over the water cells and factorised by SuperLU (scipy.sparse.linalg.splu).
Inputs are stored in full (mask, forcing, row weights, row centre weights), so the tests that read
them depend on nothing else; a hash of the generator's raw output (mask and balance) is stored too,
to notice if the generator ever changes.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from cartogenesis_amd.problem import circulation_problem  # noqa: E402

HERE = Path(__file__).resolve().parent
CASES = [(64, 0, 0.3), (128, 0, 0.3), (256, 0, 0.3), (128, 4, 0.4)]


def direct_solve(st) -> np.ndarray:
    """x = W x - f on water, 0 on land, by sparse LU in binary64."""
    across, down = st.cells_across, st.cells_down
    n = across * down
    wet = st.is_water.astype(bool)
    idx = np.arange(n).reshape(down, across)
    rows, cols, vals = [], [], []
    r = np.repeat(np.arange(down)[:, None], across, 1)
    c = np.repeat(np.arange(across)[None, :], down, 0)
    for dr, dc, w in [(0, 1, st.east_weight), (0, -1, st.west_weight), (-1, 0, st.north_weight),
                      (1, 0, st.south_weight)]:
        rr, cc = r + dr, (c + dc) % across
        ok = (rr >= 0) & (rr < down)
        src = idx[ok]
        dst = (rr * across + cc)[ok]
        keep = wet[src] & wet[dst]
        rows.append(src[keep])
        cols.append(dst[keep])
        vals.append(-w.astype(np.float64)[src[keep]])
    a = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tocsr()
    a = a + sp.eye(n, format="csr")
    wi = np.flatnonzero(wet)
    x = np.zeros(n)
    x[wi] = spl.splu(a[wi][:, wi].tocsc()).solve(-st.forcing.astype(np.float64)[wi])
    return x


def main() -> None:
    for rows, seed, land in CASES:
        p = circulation_problem(rows, seed=seed, land_fraction=land)
        st = p.levels[0]
        psi = direct_solve(st)
        digest = hashlib.sha256(p.water.tobytes() + p.balance.tobytes()).hexdigest()
        down, across = st.cells_down, st.cells_across
        roww = np.stack([st.east_weight, st.west_weight, st.north_weight, st.south_weight], 1)
        roww = roww.reshape(down, across, 4)[:, 0, :].copy()
        name = HERE / f"circulation_{rows}x{2 * rows}_seed{seed}.npz"
        np.savez_compressed(name, rows=rows, cols=2 * rows, seed=seed, land_fraction=land,
                            water=np.packbits(st.is_water.astype(bool)), forcing=st.forcing,
                            row_weights=roww.astype(np.float32),
                            row_centre=st.centre_weight.reshape(down, across)[:, 0].copy(), psi=psi,
                            generator_sha256=digest)
        print(f"wrote {name.name}: {name.stat().st_size / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
