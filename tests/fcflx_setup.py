"""A two-level composite for the fine/coarse flux points of the split solver.

A coarse block C (16 cells along the sweep direction d, 4 across) has its
cells 4..11 along d, 0..1 along the first transverse direction ORTHO1 and all
4 along ORTHO2 covered by a fine block F (16 x 4 x 8 cells of half the size;
ORTHO1, ORTHO2 = pdims(d, :), constants.F90:127-128).  Only half of C's
pencils meet the fine region, so a pencil taken for another (a transposed
face array) breaks what the composite conserves.  One stage runs as sweep() orders it on a refined level
(sweeps.F90:217-272): the fine block first, saving its boundary fluxes
(coarsebnd, save_outfluxes), then those fluxes are restricted to the coarse
faces as send_cg_coarsebnd does (fc_fluxes.F90:186-230: the 2x2 fine fluxes
summed in (ORTHO1, ORTHO2) order, times 1/2**(eff_dim-1)), and the coarse
block runs with them in place of its own fluxes at the two fine/coarse faces
(finebnd).  The fine block's ghost cells along d are the coarse parents'
values (piecewise-constant prolongation: the test's choice; any ghost values
conserve).  What is conserved: the sum over the uncovered coarse cells and
the fine cells of every component times the cell volume.
"""
import numpy as np

from oracle.pyoracle import ORTHO


NC, NB = 16, 4          # coarse cells along d; ghost layers
COV = (4, 12)           # covered coarse cells along d
COV1 = 2                # covered coarse cells along ORTHO1: 0 .. COV1-1
FIRST_C, FIRST_F = 100, 2 * (100 + COV[0])   # cg%ijkse(d, LO) of C and F


def shapes(d):
    nc = [4, 4, 4]
    nc[d] = NC
    o1, o2 = ORTHO[d]
    nf = [0, 0, 0]
    nf[d] = 2 * (COV[1] - COV[0])
    nf[o1] = 2 * COV1
    nf[o2] = 2 * nc[o2]
    dlc = [0.25, 0.25, 0.25]
    dlc[d] = 1.0 / NC
    dlf = [x / 2 for x in dlc]
    return tuple(nc), tuple(nf), tuple(dlc), tuple(dlf)


def state_at(x, y, z, gamma, magnetized):
    """Conserved variables of a smooth, non-symmetric field (cell centres)."""
    tp = 2 * np.pi
    rho = 1.0 + 0.3 * np.sin(tp * x + 0.4) * np.cos(tp * y + 0.2) + 0.1 * np.sin(tp * z + 1.1)
    vx = 0.4 * np.sin(tp * (x + 0.3 * y) + 0.7)
    vy = 0.3 * np.cos(tp * (y + 0.2 * z) + 0.1)
    vz = 0.2 * np.sin(tp * (z - 0.5 * x) + 0.5)
    p = 1.0 + 0.2 * np.cos(tp * (x - y) + 0.3)
    u = [rho, rho * vx, rho * vy, rho * vz]
    e = p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    b = psi = None
    if magnetized:
        bx = 0.5 + 0.2 * np.cos(tp * y + 0.9)
        by = 0.3 * np.sin(tp * (x + z) + 0.2)
        bz = 0.25 * np.cos(tp * x - 0.4)
        e = e + 0.5 * (bx * bx + by * by + bz * bz)
        b = np.stack([bx, by, bz])
        psi = 0.05 * np.sin(tp * (x + y + z))
    return np.stack(u + [e]), b, psi


def centres(n, nb, dl, x0, has_dir=(True, True, True)):
    g = [nb if h else 0 for h in has_dir]
    axes = [x0[d] + (np.arange(n[d] + 2 * g[d]) - g[d] + 0.5) * dl[d] for d in range(3)]
    return np.meshgrid(*axes, indexing="ij")


def fill_interior(blk, n, nb, dl, x0, fluids):
    """Every cell (ghosts too) from the smooth field, one shifted copy per
    fluid; B and psi from the magnetized fluid."""
    X, Y, Z = centres(n, nb, dl, x0)
    for f, (gam, mag) in enumerate(fluids):
        u, b, psi = state_at(X + 0.13 * f, Y, Z, gam, mag)
        blk.u[5 * f:5 * f + 5] = u
        if mag:
            blk.b[...] = b
            blk.psi[...] = psi


def prolong_ghosts(d, fine, coarse):
    """Ghost layers of F along d from the coarse parents (injection) over
    F's transverse interior, which is all a sweep along d reads."""
    _, nf, _, _ = shapes(d)
    tr = [t for t in range(3) if t != d]
    for arr_f, arr_c in ((fine.u, coarse.u), (fine.b, coarse.b), (fine.psi, coarse.psi)):
        if arr_f is None:
            continue
        lead = arr_f.ndim - 3
        pc = [np.arange(nf[t]) // 2 + NB for t in tr]   # fine interior j -> coarse parent
        for i in list(range(NB)) + list(range(NB + nf[d], nf[d] + 2 * NB)):
            ic = NB + COV[0] + (i - NB) // 2              # coarse parent along d
            sub = np.take(arr_c, ic, axis=lead + d)       # drops axis d
            sub = np.take(np.take(sub, pc[0], axis=lead + 0), pc[1], axis=lead + 1)
            idx = [slice(None)] * arr_f.ndim
            idx[lead + d] = i
            idx[lead + tr[0]] = slice(NB, NB + nf[tr[0]])
            idx[lead + tr[1]] = slice(NB, NB + nf[tr[1]])
            arr_f[tuple(idx)] = sub


def set_fine_faces(fc_f, nf, d):
    """F's whole LO and HI faces border coarse cells (coarsebnd)."""
    fc_f.index["coarse", 0][...] = FIRST_F
    fc_f.index["coarse", 1][...] = FIRST_F + nf[d] - 1


def set_coarse_faces(fc_c):
    """C's faces at 3|4 (the fine region on its HI side: ri at cell 4)
    and 11|12 (on its LO side: li at cell 11), on the pencils that meet F."""
    fc_c.index["fine", 1][:COV1, :] = FIRST_C + COV[0]
    fc_c.index["fine", 0][:COV1, :] = FIRST_C + COV[1] - 1


def restrict(fc_f, fc_c, eff_dim=3):
    """send_cg_coarsebnd + recv_cg_finebnd: fine LO face -> coarse HI side, fine HI -> coarse LO."""
    for sf, sc in ((0, 1), (1, 0)):
        for name in ("uflx", "bflx"):
            src = getattr(fc_f, name)["coarse", sf]
            if src is None:
                continue
            dst = getattr(fc_c, name)["fine", sc]
            nv, n1, n2 = src.shape
            buf = np.zeros((nv, n1 // 2, n2 // 2))
            for j in range(n1):
                for k in range(n2):
                    buf[:, j // 2, k // 2] = buf[:, j // 2, k // 2] + src[:, j, k]
            dst[:, :n1 // 2, :n2 // 2] = (1 / 2.0 ** (eff_dim - 1)) * buf


def composite_sums(d, coarse, fine):
    """Sum of every component times the cell volume over the uncovered
    coarse cells and the fine cells (interiors only)."""
    nc, nf, dlc, dlf = shapes(d)
    vc, vf = np.prod(dlc), np.prod(dlf)
    out = []
    for arr_c, arr_f in ((coarse.u, fine.u), (coarse.b, fine.b), (coarse.psi, fine.psi)):
        if arr_c is None:
            continue
        lead = arr_c.ndim - 3
        ic = [slice(NB, NB + nc[t]) for t in range(3)]
        fi = [slice(NB, NB + nf[t]) for t in range(3)]
        c = arr_c[(slice(None),) * lead + tuple(ic)]
        f = arr_f[(slice(None),) * lead + tuple(fi)]
        covered = np.zeros(nc, dtype=bool)
        box = [slice(None)] * 3
        box[d] = slice(COV[0], COV[1])
        box[ORTHO[d][0]] = slice(0, COV1)
        covered[tuple(box)] = True
        sc = c[(Ellipsis,) + tuple(np.nonzero(~covered))].sum(axis=-1) * vc
        sf = f.reshape(f.shape[:lead] + (-1,)).sum(axis=-1) * vf
        out.append(np.atleast_1d(sc + sf))
    return np.concatenate(out)


def build(po, p, cfg, d):
    """C and F with their faces marked, ghosts filled; FcFaces of both."""
    nc, nf, dlc, dlf = shapes(d)
    nvar = 5 * len(cfg.fluids)
    C = po.OBlock(nc, NB, dlc, cfg.magnetized, nvar=nvar)
    F = po.OBlock(nf, NB, dlf, cfg.magnetized, nvar=nvar)
    fill_interior(C, nc, NB, dlc, (0.0, 0.0, 0.0), cfg.fluids)
    x0 = [0.0, 0.0, 0.0]
    x0[d] = COV[0] * dlc[d]
    fill_interior(F, nf, NB, dlf, x0, cfg.fluids)
    po.fill_ghosts(p, C)
    prolong_ghosts(d, F, C)
    fc_c = po.FcFaces(nc, d, nvar, cfg.magnetized, FIRST_C)
    set_coarse_faces(fc_c)
    fc_f = po.FcFaces(nf, d, nvar, cfg.magnetized, FIRST_F)
    set_fine_faces(fc_f, nf, d)
    return C, F, fc_c, fc_f


def stage_dt(d):
    return 0.1 * shapes(d)[3][d]


def oracle_stage(po, p, d, istep, dt, ch, C, F, fc_c, fc_f, correct=True):
    """One stage of the composite: F, restriction, C, then its boundaries."""
    po.solve_cg(p, F, d, istep, dt, ch, fc=fc_f)
    restrict(fc_f, fc_c)
    po.solve_cg(p, C, d, istep, dt, ch, fc=fc_c if correct else None)
    po.fill_ghosts(p, C)
    prolong_ghosts(d, F, C)


class Absolute:
    """|u|, |b|, |psi| of a block: composite_sums of it is the scale of the sums."""

    def __init__(self, o):
        self.u = np.abs(o.u)
        self.b = np.abs(o.b) if o.b is not None else None
        self.psi = np.abs(o.psi) if o.psi is not None else None
