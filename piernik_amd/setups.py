"""Synthetic initial states for the benchmark and test problems.

Each returns a GridContainer (reference array layout, periodic ghosts
filled) plus the matching Config.  Sizes and parameters are those of the
workloads the sweep is measured on:

  sedov_hd        3D Sedov blast, adiabatic hydro, HLLC + van Leer, CFL 0.7
                  (rho=1, v=0, p=1e-5, E=1 spread over cells within 3 cells of
                  the centre; 64^3, 4 ghosts, periodic)
  mhd_turbulence  3D ideal MHD, HLLD + GLM, minmod, CFL 0.5: rho=1, p=0.6,
                  gamma=5/3, Gaussian random v and A with P(k) ~ k^-2 on
                  1<=k<=8, rms v=1, B = curl A scaled to plasma beta 1
                  (256^3, 4 ghosts, periodic)
  smooth_random   smooth seeded random state (rho, p = 1 +- 0.2, |v| <= 0.5,
                  optional B) for parity tests on small blocks
  batch_blocks    the batched block list: uneven blocks (edges drawn from
                  16, 24, 32, 48, 64), an ionized MHD fluid and a neutral
                  fluid per block, smooth seeded fields rho, p = 1 +- 0.2,
                  |v| <= 0.5 (host arrays, or generated on the GPU by
                  batch_blocks_device)
"""
import numpy as np

from . import Config, GridContainer


def fill_ghosts_periodic(cg):
    """Host-side periodic ghost refresh of every array (all directions, corners)."""
    for arr in (cg.u, cg.b, cg.psi):
        if arr is None:
            continue
        off = 1 if arr.ndim == 4 else 0
        for d in range(3):
            g, n = cg.g[d], cg.n[d]
            if g == 0:
                continue
            ax = d + off
            lo = [slice(None)] * arr.ndim
            src = [slice(None)] * arr.ndim
            lo[ax] = slice(0, g)
            src[ax] = slice(n, n + g)
            arr[tuple(lo)] = arr[tuple(src)]
            hi = [slice(None)] * arr.ndim
            src2 = [slice(None)] * arr.ndim
            hi[ax] = slice(g + n, 2 * g + n)
            src2[ax] = slice(g, 2 * g)
            arr[tuple(hi)] = arr[tuple(src2)]


def _set_state(cg, rho, vx, vy, vz, p, gamma, bx=None, by=None, bz=None):
    ui = cg.interior(cg.u)
    ui[0] = rho
    ui[1] = rho * vx
    ui[2] = rho * vy
    ui[3] = rho * vz
    e = p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    if bx is not None:
        bi = cg.interior(cg.b)
        bi[0], bi[1], bi[2] = bx, by, bz
        e = e + 0.5 * (bx * bx + by * by + bz * bz)
        cg.interior(cg.psi)[...] = 0.0
    ui[4] = e
    fill_ghosts_periodic(cg)


def sedov_hd(n=64, nb=4, gamma=5.0 / 3.0, radius_cells=3.0, e_blast=1.0, p0=1e-5):
    cg = GridContainer((n, n, n), nb=nb, dl=(1.0 / n,) * 3, magnetized=False)
    x = (np.arange(n) + 0.5) / n - 0.5
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    inside = r <= radius_cells / n
    dvol = (1.0 / n) ** 3
    rho = np.ones((n, n, n))
    zero = np.zeros((n, n, n))
    p = np.full((n, n, n), p0)
    _set_state(cg, rho, zero, zero, zero, p, gamma)
    cg.interior(cg.u)[4][inside] += e_blast / (inside.sum() * dvol)
    fill_ghosts_periodic(cg)
    cfg = Config(gamma=gamma, magnetized=False, limiter="vanleer", cfl=0.7, smalld=1e-10, smallei=1e-10)
    return cg, cfg


def gaussian_field(n, seed, kmin=1.0, kmax=8.0, slope=-2.0, ncomp=3):
    """Seeded periodic Gaussian random field(s) with shell spectrum E(k) ~ k^slope
    on kmin <= |k| <= kmax; returns (ncomp, n, n, n) with zero mean."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz1 = np.fft.rfftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k1, k1, kz1, indexing="ij")
    k = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    band = (k >= kmin) & (k <= kmax)
    amp = np.zeros_like(k)
    amp[band] = np.sqrt(k[band] ** slope / (4.0 * np.pi * k[band] ** 2))
    out = np.empty((ncomp, n, n, n))
    for c in range(ncomp):
        ph = rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)
        out[c] = np.fft.irfftn(amp * ph, s=(n, n, n), axes=(0, 1, 2))
    return out, (KX, KY, KZ, band, amp)


def mhd_turbulence(n=256, nb=4, seed=1, gamma=5.0 / 3.0, p0=0.6, rho0=1.0, vrms=1.0, beta=1.0):
    cg = GridContainer((n, n, n), nb=nb, dl=(1.0 / n,) * 3, magnetized=True)
    v, _ = gaussian_field(n, seed)
    v *= vrms / np.sqrt(np.mean(np.sum(v * v, axis=0)))
    # B = curl A, spectrally, so that div B = 0 for the continuous field
    rng_seed = seed + 1000
    a, (KX, KY, KZ, band, amp) = gaussian_field(n, rng_seed)
    ah = [np.fft.rfftn(a[c]) for c in range(3)]
    kk = (2.0 * np.pi) * np.array([KX, KY, KZ])  # box length 1
    bh = [1j * (kk[1] * ah[2] - kk[2] * ah[1]),
          1j * (kk[2] * ah[0] - kk[0] * ah[2]),
          1j * (kk[0] * ah[1] - kk[1] * ah[0])]
    b = np.array([np.fft.irfftn(bh[c], s=(n, n, n), axes=(0, 1, 2)) for c in range(3)])
    b2 = np.mean(np.sum(b * b, axis=0))
    b *= np.sqrt(2.0 * p0 / beta / b2)
    rho = np.full((n, n, n), rho0)
    p = np.full((n, n, n), p0)
    _set_state(cg, rho, v[0], v[1], v[2], p, gamma, b[0], b[1], b[2])
    cfg = Config(gamma=gamma, magnetized=True, limiter="minmod", limiter_b="minmod", cfl=0.5, cfl_glm=0.5,
                 glm_alpha=0.1, smalld=1e-4, smallei=1e-6)
    return cg, cfg


def smooth_random(n, nb=4, seed=0, magnetized=False, gamma=5.0 / 3.0, has_dir=(True, True, True), amp_b=0.7):
    """Smooth random state: rho, p in 1 +- 0.2, |v| <= 0.5 (and B of order amp_b)."""
    n = tuple(n)
    cg = GridContainer(n, nb=nb, dl=tuple(1.0 / max(1, n[d]) for d in range(3)), magnetized=magnetized,
                       has_dir=has_dir)
    rng = np.random.default_rng(seed)
    coords = np.meshgrid(*[(np.arange(n[d]) + 0.5) / n[d] for d in range(3)], indexing="ij")

    def field(lo, hi):
        f = np.zeros(n)
        for _ in range(3):
            kv = rng.integers(1, 3, size=3) * np.array([1 if h else 0 for h in has_dir])
            ph = rng.uniform(0, 2 * np.pi)
            f += np.sin(2 * np.pi * (kv[0] * coords[0] + kv[1] * coords[1] + kv[2] * coords[2]) + ph)
        f = (f - f.min()) / max(f.max() - f.min(), 1e-30)
        return lo + (hi - lo) * f

    rho = field(0.8, 1.2)
    p = field(0.8, 1.2)
    vel = [field(-0.5, 0.5) / np.sqrt(3.0) for _ in range(3)]
    if magnetized:
        bb = [field(-amp_b, amp_b) for _ in range(3)]
        _set_state(cg, rho, *vel, p, gamma, *bb)
    else:
        _set_state(cg, rho, *vel, p, gamma)
    return cg


# ---- batched block list: uneven blocks, two fluids ---------------------------

BATCH_SIZES = (16, 24, 32, 48, 64)


def batch_config(fluids=((5.0 / 3.0, True), (5.0 / 3.0, False)), limiter="vanleer", cfl=0.7, has_dir=(True, True, True)):
    """Ionized fluid (HLLD + GLM) and neutral fluid (HLLC), PIERNIK's default
    limiter (van Leer for both) and CFL 0.7, floors below any value the
    smooth state reaches."""
    return Config(fluids=list(fluids), limiter=limiter, limiter_b=limiter, cfl=cfl, cfl_glm=cfl, glm_alpha=0.1,
                  smalld=1e-4, smallei=1e-6, has_dir=has_dir)


def batch_sizes(nblocks, seed=0, sizes=BATCH_SIZES):
    """Interior edge length of each block, drawn uniformly from `sizes`."""
    rng = np.random.default_rng([seed, 7777])
    return [int(s) for s in rng.choice(np.asarray(sizes), size=nblocks)]


def _batch_params(idx, seed, nfluids, mag, has_dir):
    """Mode tables of every field of block idx: per field 3 modes of integer
    wave vectors in [-2, 2]^3 (inactive directions 0), phases and weights."""
    rng = np.random.default_rng([seed, int(idx)])
    act = np.array([1 if h else 0 for h in has_dir])
    fields = []
    for f in range(nfluids):
        modes = []
        for _ in range(5):  # rho, p, vx, vy, vz
            kv = rng.integers(-2, 3, size=(3, 3)) * act
            kv[np.all(kv == 0, axis=1)] = act  # no constant mode
            modes.append((kv, rng.uniform(0.0, 2.0 * np.pi, 3), rng.uniform(0.5, 1.0, 3)))
        fields.append(modes)
    bmodes = None
    if mag:
        # B_x(y, z), B_y(z, x), B_z(x, y): each component independent of its
        # own coordinate, so div B = 0 for the continuous field
        bmodes = []
        for c in range(3):
            kv = rng.integers(1, 3, size=3) * act
            kv[c] = 0
            bmodes.append((kv, rng.uniform(0.0, 2.0 * np.pi)))
    return fields, bmodes


def _batch_eval(xp, coords, fields, bmodes, gammas, mags, b0):
    """Conserved state from the mode tables; coords = (X, Y, Z) broadcastable
    arrays of cell centres in [0, 1).  Returns list of (5,) fluid arrays and B."""
    two_pi = 2.0 * np.pi

    def series(kv, ph, w):
        s = 0.0
        for m in range(3):
            s = s + w[m] * xp.sin(two_pi * (kv[m][0] * coords[0] + kv[m][1] * coords[1] + kv[m][2] * coords[2])
                                  + ph[m])
        return s / float(np.sum(w))

    bb = None
    if bmodes is not None:
        bb = [b0 * xp.sin(two_pi * (kv[0] * coords[0] + kv[1] * coords[1] + kv[2] * coords[2]) + ph)
              for kv, ph in bmodes]
    out = []
    for f, modes in enumerate(fields):
        rho = 1.0 + 0.2 * series(*modes[0])
        p = 1.0 + 0.2 * series(*modes[1])
        v = [(0.5 / np.sqrt(3.0)) * series(*modes[2 + d]) for d in range(3)]
        e = p / (gammas[f] - 1.0) + 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        if mags[f]:
            e = e + 0.5 * (bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2])
        out.append([rho, rho * v[0], rho * v[1], rho * v[2], e])
    return out, bb


def _shape_of(n, has_dir):
    """Block shape from an edge length (cube over the active directions) or a triple."""
    if isinstance(n, (tuple, list)):
        return tuple(int(x) for x in n)
    return tuple(int(n) if h else 1 for h in has_dir)


def _wrapped_centres(n, g):
    """Cell centres of a ghost-padded axis, ghosts at their periodic images."""
    i = np.arange(n + 2 * g) - g
    return ((i % n) + 0.5) / n


def batch_blocks(sizes, cfg, seed=0, nb=4, b0=0.3, first_id=1):
    """Host blocks (GridContainer, reference layout) of the batched list:
    rho = 1 +- 0.2, p = 1 +- 0.2, |v| <= 0.5 per fluid, B of amplitude b0 for
    the magnetized fluid, psi = 0; cubes of unit size (dl = 1/n)."""
    gammas = [g for g, _ in cfg.fluids]
    mags = [m for _, m in cfg.fluids]
    out = []
    for i, n in enumerate(sizes):
        shape = _shape_of(n, cfg.has_dir)
        cg = GridContainer(shape, nb=nb, dl=tuple(1.0 / s for s in shape), magnetized=cfg.magnetized,
                           has_dir=cfg.has_dir, cg_id=first_id + i, nfluids=cfg.nfluids)
        ax = [_wrapped_centres(shape[d], cg.g[d]) for d in range(3)]
        X, Y, Z = np.meshgrid(*ax, indexing="ij")
        fields, bmodes = _batch_params(i, seed, cfg.nfluids, cfg.magnetized, cfg.has_dir)
        fl, bb = _batch_eval(np, (X, Y, Z), fields, bmodes, gammas, mags, b0)
        for f in range(cfg.nfluids):
            for k in range(5):
                cg.u[5 * f + k] = fl[f][k]
        if cfg.magnetized:
            for c in range(3):
                cg.b[c] = bb[c]
            cg.psi[...] = 0.0
        out.append(cg)
    return out


def batch_blocks_device(sizes, cfg, seed=0, nb=4, b0=0.3, first_id=1, device="cuda"):
    """The same state generated on the GPU with torch and copied into the
    library's device mirrors (no host arrays); returns the registered
    GridContainers (their host arrays are not filled)."""
    import torch

    gammas = [g for g, _ in cfg.fluids]
    mags = [m for _, m in cfg.fluids]
    out = []
    for i, n in enumerate(sizes):
        shape = _shape_of(n, cfg.has_dir)
        cg = _DeviceBlock(shape, nb, cfg, first_id + i)
        ax = [torch.as_tensor(_wrapped_centres(shape[d], cg.g[d]), dtype=torch.float64, device=device)
              for d in range(3)]
        X = ax[0].view(1, 1, -1)
        Y = ax[1].view(1, -1, 1)
        Z = ax[2].view(-1, 1, 1)
        fields, bmodes = _batch_params(i, seed, cfg.nfluids, cfg.magnetized, cfg.has_dir)
        fl, bb = _batch_eval(torch, (X, Y, Z), fields, bmodes, gammas, mags, b0)
        full = (cg.ng[2], cg.ng[1], cg.ng[0])
        u = torch.stack([torch.broadcast_to(torch.as_tensor(c, dtype=torch.float64, device=device), full)
                         for f in range(cfg.nfluids) for c in fl[f]], dim=-1).contiguous()
        b = psi = None
        if cfg.magnetized:
            b = torch.stack([torch.broadcast_to(c, full) for c in bb], dim=-1).contiguous()
            psi = torch.zeros(full, dtype=torch.float64, device=device)
        cg.upload_device(u, b, psi)
        del u, b, psi
        out.append(cg)
    return out


class _DeviceBlock(GridContainer):
    """A GridContainer registered without host arrays (device-resident)."""

    def __init__(self, shape, nb, cfg, cg_id):
        self.n = tuple(shape)
        self.nb = nb
        self.has_dir = tuple(cfg.has_dir)
        self.g = tuple(nb if h else 0 for h in self.has_dir)
        self.ng = tuple(self.n[d] + 2 * self.g[d] for d in range(3))
        self.dl = tuple(1.0 / s for s in self.n)
        self.magnetized = cfg.magnetized
        self.nfluids = cfg.nfluids
        self.nvar = 5 * self.nfluids
        self.u = self.uh = self.b = self.bh = self.psi = self.psih = None
        self.cg_id = int(cg_id)
        self.registered = False


# ---- differentiable hydro rollout ------------------------------------------------


def decompose(cg, cuts, first_id=None):
    """Split the periodic domain of `cg` into blocks at the interior cut
    positions cuts[d] (sorted, 0 < c < n[d]; ignored in inactive
    directions; every block at least nb cells wide, so a ghost layer never
    reaches past the adjoining block), blocks numbered x fastest.  Each block has cg's cell size
    and nb, and its state (ghost cells included) is cg's interior with
    periodic wrap.  Returns (blocks, nbr, boxes): GridContainers, nbr[i][d][s]
    the index of the block whose interior adjoins block i's ghost layers on
    side s (0 low, 1 high) of direction d (periodic across the domain), and
    boxes[i] = (offset, size) in cells."""
    edges = []
    for d in range(3):
        c = sorted(int(x) for x in cuts[d]) if cg.has_dir[d] else []
        if any(x <= 0 or x >= cg.n[d] for x in c):
            raise ValueError(f"cut outside 1..{cg.n[d] - 1} along {'xyz'[d]}")
        edges.append([0] + c + [cg.n[d]])
        if cg.has_dir[d] and min(np.diff(edges[-1])) < cg.nb:
            raise ValueError(f"blocks along {'xyz'[d]} must be at least nb = {cg.nb} cells wide")
    cnt = [len(e) - 1 for e in edges]
    arrays = [("u", cg.u), ("b", cg.b), ("psi", cg.psi)]
    blocks, boxes, nbr = [], [], []
    for k in range(cnt[2]):
        for j in range(cnt[1]):
            for i in range(cnt[0]):
                ijk = (i, j, k)
                off = tuple(edges[d][ijk[d]] for d in range(3))
                size = tuple(edges[d][ijk[d] + 1] - off[d] for d in range(3))
                kw = {} if first_id is None else {"cg_id": first_id + len(blocks)}
                blk = GridContainer(size, nb=cg.nb, dl=cg.dl, magnetized=cg.magnetized, has_dir=cg.has_dir,
                                    nfluids=cg.nfluids, **kw)
                idx = [(np.arange(-blk.g[d], size[d] + blk.g[d]) + off[d]) % cg.n[d] + cg.g[d] for d in range(3)]
                for name, a in arrays:
                    if a is None:
                        continue
                    dst = getattr(blk, name)
                    if a.ndim == 4:
                        dst[...] = a[:, idx[0]][:, :, idx[1]][:, :, :, idx[2]]
                    else:
                        dst[...] = a[idx[0]][:, idx[1]][:, :, idx[2]]
                row = []
                for d in range(3):
                    lo, hi = list(ijk), list(ijk)
                    lo[d] = (ijk[d] - 1) % cnt[d]
                    hi[d] = (ijk[d] + 1) % cnt[d]
                    row.append([lo[0] + cnt[0] * (lo[1] + cnt[1] * lo[2]), hi[0] + cnt[0] * (hi[1] + cnt[1] * hi[2])])
                nbr.append(row)
                blocks.append(blk)
                boxes.append((off, size))
    return blocks, nbr, boxes


def assemble(blocks, boxes, name, n):
    """The interior of array `name` ("u", "b" or "psi") over the whole domain
    of n cells from the blocks of decompose()."""
    first = getattr(blocks[0], name)
    out = np.zeros(first.shape[:-3] + tuple(n), dtype=first.dtype) if first.ndim == 4 else np.zeros(tuple(n))
    for blk, (off, size) in zip(blocks, boxes):
        sl = tuple(slice(off[d], off[d] + size[d]) for d in range(3))
        a = blk.interior(getattr(blk, name))
        if a.ndim == 4:
            out[(slice(None),) + sl] = a
        else:
            out[sl] = a
    return out


# problems/cpaw/problem.par: the wave the reference's CPAW test measures
CPAW_PAR = dict(d0=1.0, p0=0.1, vx0=0.0, vy0=0.0, vz0=0.0, A=0.1, vA=1.0, kx=6.283185307, ky=12.566370614, kz=0.0)
CPAW_TEND = 0.447213595
# L2 error norm of B after one period, problems/cpaw/problem.par:2-13
CPAW_L2 = {(32, 16): 0.00996214, (48, 24): 0.00559514, (64, 32): 0.00347388, (96, 48): 0.00171573,
           (128, 64): 0.00102726, (192, 96): 0.00049221, (256, 128): 0.00029084}


def cpaw(nx=128, ny=64, nb=4, gamma=5.0 / 3.0, **par):
    """Circularly polarised Alfven wave (problems/cpaw/initproblem.F90:137-205,
    problem.par): 2D periodic domain [0, 1] x [0, 0.5] of nx x ny cells,
    the wave vector (kx, ky, 0) rotated from the x axis, v' = (vx0, vy0 +
    A sin phi, vz0 + A cos phi), B' = (vA sqrt(d0), -sqrt(d0) A sin phi,
    -sqrt(d0) A cos phi), p = p0, every cell ghosts included from the
    formula.  Returns (GridContainer, Config, ini_B) with ini_B the initial
    interior B (3, nx, ny, 1) the L2 error is measured against; the unsplit
    solver's settings: HLLD + GLM, van Leer, CFL 0.3, cfl_glm 0.3,
    smalld = smallei = 1e-5, ionized gamma 5/3."""
    q = dict(CPAW_PAR)
    q.update(par)
    d0, p0, A, vA = q["d0"], q["p0"], q["A"], q["vA"]
    kx, ky, kz = q["kx"], q["ky"], q["kz"]
    hd = (True, True, False)
    dl = (1.0 / nx, 0.5 / ny, 1.0)
    cg = GridContainer((nx, ny, 1), nb=nb, dl=dl, magnetized=True, has_dir=hd)
    x = (np.arange(cg.ng[0]) - nb + 0.5) * dl[0]
    y = (np.arange(cg.ng[1]) - nb + 0.5) * dl[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    Z = np.full_like(X, 0.5)
    aa = np.arctan(ky / kx)
    g = np.arctan(np.cos(aa) * (kz / kx))
    kp = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    rot = np.array([[np.cos(aa) * np.cos(g), -np.sin(aa), -np.cos(aa) * np.sin(g)],
                    [np.sin(aa) * np.cos(g), np.cos(aa), -np.sin(aa) * np.sin(g)],
                    [np.sin(g), 0.0, np.cos(g)]])
    phi = kp * ((rot[0, 0] * X + rot[1, 0] * Y) + rot[2, 0] * Z)
    vp = [q["vx0"] + 0.0 * phi, q["vy0"] + A * np.sin(phi), q["vz0"] + A * np.cos(phi)]
    bp = [vA * np.sqrt(d0) + 0.0 * phi, -np.sqrt(d0) * A * np.sin(phi), -np.sqrt(d0) * A * np.cos(phi)]
    v = [(rot[i, 0] * vp[0] + rot[i, 1] * vp[1]) + rot[i, 2] * vp[2] for i in range(3)]
    b = [(rot[i, 0] * bp[0] + rot[i, 1] * bp[1]) + rot[i, 2] * bp[2] for i in range(3)]
    cg.u[0, ..., 0] = d0
    for i in range(3):
        cg.u[1 + i, ..., 0] = v[i]
        cg.b[i, ..., 0] = b[i]
    ekin = 0.5 * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) * d0
    emag = 0.5 * ((b[0] * b[0] + b[1] * b[1]) + b[2] * b[2])
    cg.u[4, ..., 0] = (p0 / (gamma - 1.0) + ekin) + emag
    cfg = Config(gamma=gamma, magnetized=True, limiter="vanleer", cfl=0.3, cfl_glm=0.3, smalld=1e-5, smallei=1e-5,
                 has_dir=hd)
    return cg, cfg, np.array(cg.interior(cg.b))


def cpaw_l2(b, ini_b):
    """The CPAW error norm (problems/cpaw/initproblem.F90:268-282): over the
    active directions (x, y), sqrt(sum (B_ini - B)^2 / sum B_ini^2)."""
    num = sum(float(np.sum((ini_b[d] - b[d]) ** 2)) for d in range(2))
    den = sum(float(np.sum(ini_b[d] ** 2)) for d in range(2))
    return float(np.sqrt(num / den))


def unit_gaussian(n, seed, kmax=6.0):
    """Seeded periodic Gaussian random field on n^3 with 0 < |k| <= kmax,
    normalised to zero mean and unit variance."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz = np.fft.rfftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k1, k1, kz, indexing="ij")
    k = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    band = (k > 0) & (k <= kmax)
    ph = rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)
    f = np.fft.irfftn(ph * band, s=(n, n, n), axes=(0, 1, 2))
    return f / np.sqrt(np.mean(f * f))


def adjoint_hydro(n=128, seed=0, gamma=5.0 / 3.0, amp=0.1):
    """Initial state and target of the adjoint rollout: rho = 1 + amp*G
    (G unit-variance Gaussian, |k| <= 6), p = 1, v = 0; target density from
    seed + 1.  Returns (u (5, n, n, n) interior, rho_target (n, n, n), cfg)."""
    rho = 1.0 + amp * unit_gaussian(n, seed)
    u = np.zeros((5, n, n, n))
    u[0] = rho
    u[4] = 1.0 / (gamma - 1.0)
    target = 1.0 + amp * unit_gaussian(n, seed + 1)
    cfg = Config(gamma=gamma, magnetized=False, limiter="vanleer", cfl=0.5, smalld=1e-10, smallei=1e-10)
    return u, target, cfg


def to_device_layout(u, nb=4, has_dir=(True, True, True)):
    """(C, nx, ny, nz) interior array (numpy or torch; (nx, ny, nz) for one
    component) -> torch tensor (nzg, nyg, nxg, C) on the GPU with periodic
    ghost cells in the active directions: the reference layout
    u(C, nxg, nyg, nzg) (or (nzg, nyg, nxg) for one component)."""
    import torch

    t = torch.as_tensor(np.ascontiguousarray(u) if isinstance(u, np.ndarray) else u, dtype=torch.float64)
    scalar = t.dim() == 3
    if scalar:
        t = t.unsqueeze(0)
    for d, ax in enumerate((1, 2, 3)):
        if has_dir[d]:
            t = torch.cat([t.narrow(ax, t.shape[ax] - nb, nb), t, t.narrow(ax, 0, nb)], dim=ax)
    t = t.permute(3, 2, 1, 0).contiguous()
    if scalar:
        t = t[..., 0].contiguous()
    return t.cuda()


def interior_of(t, nb=4, has_dir=(True, True, True)):
    """Inverse view of to_device_layout: (nzg, nyg, nxg[, C]) -> (C, nx, ny, nz) or (nx, ny, nz)."""
    g = [nb if h else 0 for h in has_dir]
    sl = tuple(slice(g[d], t.shape[2 - d] - g[d]) for d in (2, 1, 0))
    v = t[sl]
    return v.permute(3, 2, 1, 0) if v.dim() == 4 else v.permute(2, 1, 0)


def orszag_tang_2d(n, seed=0, amp=1.0, perturb=0.0, gamma=5.0 / 3.0):
    """Orszag-Tang vortex on the unit square (2D, n x n): rho = 25/(36 pi),
    p = 5/(12 pi), v = amp*(-sin 2pi y, sin 2pi x), B = amp*B0*(-sin 2pi y,
    sin 4pi x), B0 = 1/sqrt(4 pi), psi = 0; `perturb` adds seeded smooth
    perturbations to every component, psi included (a generic state).  Returns interior
    arrays u (5, n, n, 1), b (3, n, n, 1), psi (n, n, 1) and the Config."""
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    rng = np.random.default_rng(seed)

    def smooth():
        f = np.zeros((n, n))
        for _ in range(3):
            kx, ky = rng.integers(1, 3, 2)
            f += np.sin(2 * np.pi * (kx * X + ky * Y) + rng.uniform(0, 2 * np.pi))
        return f / 3.0

    rho = 25.0 / (36.0 * np.pi) * (1.0 + perturb * smooth())
    p = 5.0 / (12.0 * np.pi) * (1.0 + perturb * smooth())
    vx, vy = -amp * np.sin(2 * np.pi * Y), amp * np.sin(2 * np.pi * X)
    b0 = 1.0 / np.sqrt(4.0 * np.pi)
    bx, by = -amp * b0 * np.sin(2 * np.pi * Y), amp * b0 * np.sin(4 * np.pi * X)
    # a generic state: the plain vortex has v_x, B_x constant along x and
    # v_y, B_y constant along y (limited slopes exactly 0)
    vx = vx + perturb * smooth()
    vy = vy + perturb * smooth()
    bx = bx + perturb * b0 * smooth()
    by = by + perturb * b0 * smooth()
    vz = perturb * smooth()
    bz = perturb * b0 * smooth()
    u = np.zeros((5, n, n, 1))
    u[0, ..., 0] = rho
    u[1, ..., 0] = rho * vx
    u[2, ..., 0] = rho * vy
    u[3, ..., 0] = rho * vz
    u[4, ..., 0] = (p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
                    + 0.5 * (bx * bx + by * by + bz * bz))
    b = np.zeros((3, n, n, 1))
    b[0, ..., 0] = bx
    b[1, ..., 0] = by
    b[2, ..., 0] = bz
    psi = np.zeros((n, n, 1))
    psi[..., 0] = perturb * b0 * smooth()
    cfg = Config(gamma=gamma, magnetized=True, limiter="vanleer", cfl=0.5, cfl_glm=0.5, smalld=1e-10,
                 smallei=1e-10, has_dir=(True, True, False))
    return u, b, psi, cfg


OT_SIZES = (64, 96, 128, 192, 256)


def orszag_tang_ensemble(ninst=256, seed=0, sizes=OT_SIZES, spread=0.1):
    """Instances of the 2D Orszag-Tang vortex with sizes drawn from `sizes`
    and per-instance amplitudes 1 +- spread (seeded).  Returns a list of
    (n, u, b, psi) interior arrays and the Config."""
    rng = np.random.default_rng([seed, 5555])
    out = []
    cfg = None
    for i in range(ninst):
        n = int(rng.choice(np.asarray(sizes)))
        amp = 1.0 + spread * rng.uniform(-1.0, 1.0)
        u, b, psi, cfg = orszag_tang_2d(n, seed=i, amp=amp)
        out.append((n, u, b, psi))
    return out, cfg
