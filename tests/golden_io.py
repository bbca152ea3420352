"""Loads the golden fixtures of tests/golden/ (written by tests/golden/make_golden.py)."""
import glob
import json
import os

import numpy as np
import scipy.sparse as sp

from orctm_amd import synth

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(HERE, "*.npz")))


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    fl = json.loads(str(z["flags"]))
    spec = synth.GridSpec(fl["nx"], fl["ny"], fl["nz"], cyclic_x=fl["cyclic_x"], cyclic_y=fl["cyclic_y"],
                          open_w=fl["open_w"], open_e=fl["open_e"], open_n=fl["open_n"], open_s=fl["open_s"])
    g = synth.Grid(spec, z["dzw"], z["dlxp"], z["dlyp"], z["dlxu"], z["dlyv"], z["ddpo"], z["weto"])
    N = g.cells
    M = sp.csr_matrix((z["A_val"], (z["A_row"], z["A_col"])), shape=(N, N))
    return g, np.asfortranarray(z["divg"]), M, z["b"], z["p_direct"]
