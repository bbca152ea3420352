"""Small seeded systems shared by the CPU and GPU tests (each solvable by the oracle in well under a second).

They cover the boundary cases of compute_matrix (mo_petsc.f90:543-782): open boundaries (wet halo,
Dirichlet 0), closed boundaries (land halo, Neumann), periodic x and y, random land with partial bottom
cells, a non-uniform DLXP profile, a single layer, and pitch padding (nx not a multiple of 32).
"""
import numpy as np

from orctm_amd import synth

SPECS = {
    # regional tile, open boundaries all round, 25 % land, partial cells in (0.1, 1]
    "regional": synth.GridSpec(70, 37, 13, smooth=4.0, seed=11),
    # ISW transect: periodic y, closed x ends, Gaussian ridge
    "transect": synth.GridSpec(96, 4, 40, dx=50.0, dy=50.0, dz0=2.0, dz1=40.0, depth=1500.0, bathy="ridge",
                               ridge_height=800.0, ridge_width=500.0, cyclic_y=True, open_w=False, open_e=False,
                               open_n=False, open_s=False, frac_lo=0.2, sigma_u=0.1, seed=12),
    # periodic x, closed north/south
    "cyclic_x": synth.GridSpec(64, 21, 9, cyclic_x=True, open_n=False, open_s=False, smooth=3.0, seed=13),
    # closed basin, odd sizes
    "closed": synth.GridSpec(33, 29, 17, open_w=False, open_e=False, open_n=False, open_s=False, smooth=3.0,
                             land_frac=0.4, seed=14),
    # doubly periodic, flat bottom
    "cyclic_xy": synth.GridSpec(40, 24, 7, bathy="flat", depth=900.0, cyclic_x=True, cyclic_y=True, seed=15),
    # stretched sponge-like DLXP profile (as the Sect.3.5 ddxRef files stretch the ends)
    "stretched_dx": synth.GridSpec(48, 10, 12, dx_profile=np.concatenate([np.geomspace(2600.0, 300.0, 8),
                                                                          np.full(32, 200.0),
                                                                          np.geomspace(300.0, 2600.0, 8)]),
                                   open_n=False, open_s=False, smooth=2.0, seed=16),
    # one layer only (surface Dirichlet row with no bottom neighbour)
    "one_layer": synth.GridSpec(20, 12, 1, depth=50.0, dz0=50.0, dz1=50.0, bathy="flat", seed=17),
}


def build(name):
    spec = SPECS[name]
    return synth.build(spec)
