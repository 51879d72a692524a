"""States on which care_for_positives acts (sources.F90:299-433): the
density floor, the internal-energy floor and the negative-value flags.

Shared by the oracle pins and the GPU floor tests.  Smooth random states
(rho, p = 1 +- 0.2) with smalld = 0.9, so every cell below 0.9 is raised in
every stage, and smallei above the smallest internal energies; three cells
have their total energy set to zero, so their internal energy is negative
(-e_kin - e_mag) and the ei_negative flag is raised where disallow_negatives
is set.
"""
import numpy as np

import piernik_amd as pk
from piernik_amd import setups

ZERO_E_CELLS = [(15, 3, 6), (2, 9, 1), (9, 5, 8)]
NEG_RHO_CELLS = [(3, 2, 2), (11, 7, 5), (18, 4, 8), (6, 10, 3)]
EPS = np.finfo(np.float64).eps


def floor_block(shape=(20, 12, 10), magnetized=False, seed=0, neg_rho=False, fluids=None):
    """Block and Config (at least 19 x 11 x 9 cells, for the cells set
    below); neg_rho: also four isolated cells at rest with density -1e-4
    (hydro only: their periodic images at the pencil ends stay negative
    through the first stage, which is what raises dn_negative)."""
    cg = setups.smooth_random(shape, seed=seed, magnetized=magnetized)
    ui = cg.interior(cg.u)
    for (i, j, k) in ZERO_E_CELLS:
        ui[4, i, j, k] = 0.0
    if neg_rho:
        for (i, j, k) in NEG_RHO_CELLS:
            ui[0, i, j, k] = -1e-4
            ui[1:4, i, j, k] = 0.0
    setups.fill_ghosts_periodic(cg)
    cfg = pk.Config(gamma=5.0 / 3.0, magnetized=magnetized, limiter="vanleer", cfl=0.7,
                    smalld=1e-6 if neg_rho else 0.9, smallei=1.5)
    return cg, cfg


def mass_tolerance(cg, stage_sweeps):
    """Bound on |GPU - reference form| of the magic mass.  The reference
    adds and removes each pencil's interior mass S*dvol around the floor
    (sources.F90:356-358), so every pencil and stage rounds at the scale of
    that mass: summed, a few eps times the block's mass per stage sweep.
    The GPU sums the added mass exactly and rounds once."""
    rho = np.abs(cg.interior(cg.u)[0]).sum()
    dvol = float(np.prod([cg.dl[d] for d in range(3) if cg.has_dir[d]]))
    return 8.0 * EPS * stage_sweeps * (rho * dvol + 1.0)
