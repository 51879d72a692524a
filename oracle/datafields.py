"""Output fields of PIERNIK's data files, restated in numpy.

TEST INFRASTRUCTURE ONLY: the checker of pk_datafield (libpiernik_hip) and of
the GDF files written from it.  Only tests/ may import it.

datafields_hdf5 (src/IO/hdf5/data_hdf5.F90:496-580) over the interior cells
(RNG) of a block in the reference layout u(nvar, nxg, nyg, nzg),
b(3, ...), psi(...), with ekin, emag and sq_sum3 of src/base/func.F90:55-88:
    sq_sum3(a, b, c) = a**2 + b**2 + c**2      (left to right)
    emag(b)          = half * sq_sum3(b)
    ekin(m, dn)      = emag(m) / dn
prei and ethi subtract emag_c (cell-centred B, cc_mag) from the ionized
fluid's internal energy; psi is the qna array get_data_from_cg falls back to
(:1021-1026).  Every operation is one IEEE operation on float64 arrays in the
reference's order (numpy does not contract), so the values are the
reference's to the last bit.

Fluid letters (common_shortcuts, common_hdf5.F90:418-542): 'i' the ionized
fluid -- the magnetized one, else the first; 'n' the neutral one -- the other
of two fluids, or the only fluid of a run without magnetic field.
"""
import numpy as np


def sq_sum3(a, b, c):
    return a * a + b * b + c * c


def emag(bx, by, bz):
    return 0.5 * sq_sum3(bx, by, bz)


def ekin(mx, my, mz, dn):
    return emag(mx, my, mz) / dn


def _fluid(letter, fluids):
    """fluids: [(gamma, magnetized), ...] in flind order -> (index, gamma)."""
    mag = [i for i, (_, m) in enumerate(fluids) if m]
    if letter == "i":
        f = mag[0] if mag else 0
    elif letter == "n":
        if len(fluids) == 2:
            f = 1 - mag[0] if mag else 1
        elif mag:
            raise KeyError("no neutral fluid")
        else:
            f = 0
    else:
        raise KeyError(letter)
    return f, fluids[f][0]


def datafield(name, u, b, psi, fluids, nb, has_dir=(True, True, True)):
    """Field `name` over the interior cells, shape (nx, ny, nz)."""
    g = [nb if h else 0 for h in has_dir]
    rng = tuple(slice(g[d], u.shape[1 + d] - g[d]) for d in range(3))
    U = lambda c: u[(c,) + rng]
    B = (lambda c: b[(c,) + rng]) if b is not None else None
    emag_c = emag(B(0), B(1), B(2)) if B is not None else None
    if name == "psi":
        return psi[rng].copy()
    if name in ("magx", "magy", "magz"):
        return B("xyz".index(name[3])).copy()
    if name == "magB":
        return np.sqrt(2.0 * emag_c)
    if name == "v":
        f, _ = _fluid("i" if (len(fluids) == 2 or b is not None) else "n", fluids)
        o = 5 * f
        return np.sqrt(sq_sum3(U(o + 1), U(o + 2), U(o + 3))) / U(o)
    kind, letter = name[:-1], name[-1]
    f, gam = _fluid(letter, fluids)
    o = 5 * f
    if kind == "den":
        return U(o).copy()
    if kind == "ene":
        return U(o + 4).copy()
    if kind in ("vlx", "vly", "vlz"):
        return U(o + 1 + "xyz".index(kind[2])) / U(o)
    if kind in ("momx", "momy", "momz"):
        return U(o + 1 + "xyz".index(kind[3])).copy()
    if kind in ("pre", "eth"):
        eth = U(o + 4) - ekin(U(o + 1), U(o + 2), U(o + 3), U(o))
        eth = eth - (emag_c if (letter == "i" and emag_c is not None) else 0.0)
        return (gam - 1.0) * eth if kind == "pre" else eth / U(o)
    raise KeyError(name)
