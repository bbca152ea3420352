"""Writes the golden fixtures tests/golden/*.npz (run from the repository root):

    python tests/golden/make_golden.py

Each fixture holds one small system's inputs in the reference's Fortran layout (DZW, DLXP, DLYP, DLXU,
DLYV, DDPO, WETO, DIVG and the boundary flags), the operator written by the dense transliteration of
compute_matrix in tests/test_oracle_operator.py (not by the oracle), and the FP64 direct solution of
that operator with b = DIVG*WETO (SciPy SuperLU). The reference ships no golden vectors of its own;
these pin the oracle's assembly and solve independently of the C restatement.
"""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from orctm_amd import synth  # noqa: E402
from test_oracle_operator import dense_reference  # noqa: E402

FIXTURES = {
    "golden_regional": synth.GridSpec(20, 12, 8, smooth=2.0, dmax=2000.0, depth=2000.0, seed=101),
    "golden_transect": synth.GridSpec(24, 4, 10, dx=50.0, dy=50.0, dz0=2.0, dz1=40.0, depth=300.0, bathy="ridge",
                                      ridge_height=150.0, ridge_width=300.0, cyclic_y=True, open_w=False,
                                      open_e=False, open_n=False, open_s=False, frac_lo=0.2, sigma_u=0.1, seed=102),
    "golden_cyclic_x": synth.GridSpec(16, 9, 6, cyclic_x=True, open_n=False, open_s=True, smooth=2.0, seed=103),
}


def main():
    for name, spec in FIXTURES.items():
        g, divg = synth.build(spec)
        M = dense_reference(g)
        b = g.interior(divg) * g.interior(g.weto)
        p = spla.spsolve(sp.csc_matrix(M), b)
        Mc = sp.coo_matrix(M)
        flags = dict(cyclic_x=spec.cyclic_x, cyclic_y=spec.cyclic_y, open_w=spec.open_w, open_e=spec.open_e,
                     open_n=spec.open_n, open_s=spec.open_s, nx=spec.nx, ny=spec.ny, nz=spec.nz)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), dzw=g.dzw, dlxp=g.dlxp, dlyp=g.dlyp, dlxu=g.dlxu,
                            dlyv=g.dlyv, ddpo=g.ddpo, weto=g.weto, divg=divg, A_row=Mc.row, A_col=Mc.col,
                            A_val=Mc.data, b=b, p_direct=p, flags=json.dumps(flags))
        r = np.linalg.norm(b - M @ p) / np.linalg.norm(b)
        print(f"{name}: {g.cells} cells, direct rel.res {r:.2e}")


if __name__ == "__main__":
    main()
