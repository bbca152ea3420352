"""Input and output files of the Fortran host of the binding (tests/fortran/driver.F90): the stream
layout its header comment states, written from and read into Fortran-ordered arrays."""
import os

import numpy as np

from cases import build
from orctm_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
FBUILD = os.path.join(HERE, "fortran", "_build")


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def inputs():
    g, divg = build("regional")
    assert g.KE == 13 and not g.spec.cyclic_x and not g.spec.cyclic_y  # what the executables are built for
    v = synth.make_velocity(g)
    b = synth.make_barotropic(g)
    b1o = np.asfortranarray(synth.normal(91, 1, b.IE * b.JE).reshape(b.JE, b.IE).T)
    return g, divg, v, b, b1o


def write_input(path, g, divg, v, b, b1o):
    sp = g.spec
    hdr = np.array([g.IE, g.JE, 0, 0, sp.open_w, sp.open_e, sp.open_n, sp.open_s], dtype=np.int32)
    sc = np.array([v.dt, v.dti, b.g, b.conn, b.cono, b.stabn, b.sorpar], dtype=np.float64)
    arrays = [g.dlxp, g.dlyp, g.dlxu, g.dlyv, g.ddpo, g.weto, g.dzw, divg,
              v.amsuo, v.dduo, v.dtdxuo, v.dlyu, v.amsue, v.ddue, v.dpye, v.dlxv, v.dtdxpo, v.dtdyo,
              v.uoo, v.uoo_1, v.voe, v.voe_1, v.wno, v.wno_1, v.zo,
              b.deuto, b.deute, b.u1o, b.uzo, b.v1e, b.vze, b1o]
    with open(path, "wb") as f:
        hdr.tofile(f)
        sc.tofile(f)
        for a in arrays:
            np.asarray(a, dtype=np.float64).ravel(order="F").tofile(f)


def read_output(path, shapes):
    raw = np.fromfile(path, dtype=np.float64)
    out, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(raw[o:o + n].reshape(s, order="F"))
        o += n
    assert o == raw.size
    return out
