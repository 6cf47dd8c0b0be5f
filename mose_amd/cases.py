"""Case builders for the explicit residual/update path.

Each builder returns a `Case`: configuration, block nodes (blk%node layout), the bc.txt
records in file order (sorted by block, then face, as Read_BCfile expects,
src/lib/io/IO_BC.f90:136-414), and the primitive initial state of every block in the
blk%P layout.  The same Case feeds the HIP library and the CPU oracle.

The workloads are the five fixed configurations of this path:
  * Sod shock tube, 1000 cells (the reference's own test/1D/Sod79 case family),
  * isentropic vortex on a periodic 256^3 block,
  * 64 seeded Cartesian blocks joined by 1-to-1 connections,
  * a 2-D supersonic channel with a circular-arc bump,
  * batched Taylor-Green boxes.
"""
from dataclasses import dataclass, field

import numpy as np

from . import abi

GAMMA = 1.4
RGAS_NONDIM = 1.0 / 1.4      # test/common/nondim-Air: R = Runiv/11640.3086 ~ 1/1.4


@dataclass
class Block:
    dims: tuple
    nodes: np.ndarray            # (nk+1, nj+1, ni+1, 3) == Fortran (3, 0:ni, 0:nj, 0:nk)
    P: np.ndarray = None         # (nk+4, nj+4, ni+4, 5) == Fortran (5, -1:ni+2, ...)


@dataclass
class Case:
    name: str
    cfg: abi.MoseCfg
    blocks: list
    bc: np.ndarray
    t_end: float = None
    nsteps: int = None
    meta: dict = field(default_factory=dict)

    @property
    def ncells(self):
        return int(sum(np.prod(b.dims) for b in self.blocks))


# ----------------------------------------------------------------------------- grids
def cartesian_nodes(dims, lo, hi):
    """Uniform Cartesian node array for a block of `dims` cells over the box [lo, hi]."""
    ni, nj, nk = dims
    x = np.linspace(lo[0], hi[0], ni + 1)
    y = np.linspace(lo[1], hi[1], nj + 1)
    z = np.linspace(lo[2], hi[2], nk + 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.ascontiguousarray(np.stack([X, Y, Z], axis=-1))


def cell_centres(nodes):
    """Cell centres (nk, nj, ni, 3) as the mean of the 8 cell nodes."""
    n = nodes
    return 0.125 * (n[:-1, :-1, :-1] + n[:-1, :-1, 1:] + n[:-1, 1:, :-1] + n[:-1, 1:, 1:] +
                    n[1:, :-1, :-1] + n[1:, :-1, 1:] + n[1:, 1:, :-1] + n[1:, 1:, 1:])


def padded_state(dims, interior_prim):
    """blk%P array with the interior set from (nk, nj, ni, 5) and ghosts filled by copy
    of the nearest interior cell (ghosts are refilled by the solver before use)."""
    g = abi.GC
    P = np.pad(np.asarray(interior_prim, dtype=np.float64), ((g, g), (g, g), (g, g), (0, 0)), mode="edge")
    assert P.shape == abi.padded_shape(dims)
    return np.ascontiguousarray(P)


# ------------------------------------------------------------------------- bc records
def face_cells(dims, f):
    """(i, j, k) 1-based boundary cells of face f (1/2 i-min/max, 3/4 j, 5/6 k) as an (n, 3)
    int array, ordered with the first tangential index fastest."""
    ni, nj, nk = dims
    d = (f - 1) // 2
    n = [ni, nj, nk]
    t = [a for a in range(3) if a != d]
    A, B = np.meshgrid(np.arange(1, n[t[0]] + 1), np.arange(1, n[t[1]] + 1), indexing="xy")
    out = np.empty((A.size, 3), dtype=np.int32)
    out[:, d] = 1 if f % 2 == 1 else n[d]
    out[:, t[0]] = A.ravel()
    out[:, t[1]] = B.ravel()
    return out


def bc_records(block_faces):
    """Build the bc.txt record array.

    block_faces: list over blocks of a list of 6 face specs.  A face spec is either an
    int BC type (0, 300, 400), a dict {"type": 405, mach, T, p, alpha, beta, rel_fac},
    {"type": 401, T (the stagnation T0), p (p0), alpha, beta, rel_fac}, {"type": 406, p (the
    back pressure), rel_fac},
    or a dict {"type": 101|201, "bs": donor block, "fs": donor face} for an aligned
    1-to-1 connection (orientation d11,d12,d21,d22 = 1,0,0,1)."""
    parts = []
    dims_of = [bf["dims"] for bf in block_faces]
    for b, bf in enumerate(block_faces, start=1):
        dims = bf["dims"]
        for f in range(1, 7):
            spec = bf["faces"][f - 1]
            if not isinstance(spec, dict):
                spec = {"type": int(spec)}
            cells = face_cells(dims, f)
            arr = np.zeros(len(cells), dtype=abi.BC_DTYPE)
            arr["b"], arr["f"], arr["type"] = b, f, spec["type"]
            arr["i"], arr["j"], arr["k"] = cells[:, 0], cells[:, 1], cells[:, 2]
            if spec["type"] in (101, 201):
                bs, fs = spec["bs"], spec["fs"]
                sd = dims_of[bs - 1]
                di = (f - 1) // 2
                if (fs - 1) // 2 != di:
                    raise ValueError("only same-direction connections are generated")
                src = cells.copy()
                src[:, di] = 1 if fs % 2 == 1 else sd[di]
                arr["bs"], arr["fs"] = bs, fs
                arr["is"], arr["js"], arr["ks"] = src[:, 0], src[:, 1], src[:, 2]
                arr["d11"], arr["d22"] = 1, 1
            elif spec["type"] == 405:
                for key in ("mach", "T", "p", "alpha", "beta", "rel_fac"):
                    arr[key] = spec[key]
            elif spec["type"] == 401:
                for key in ("T", "p", "alpha", "beta", "rel_fac"):
                    arr[key] = spec[key]
            elif spec["type"] == 406:
                arr["p"], arr["rel_fac"] = spec["p"], spec.get("rel_fac", 1.0)
            parts.append(arr)
    return np.concatenate(parts)


def expected_bc_count(dims_list):
    """Sum_b 2 (nj nk + ni nk + ni nj): the record count Check_BC requires (IO_BC.f90:62-71)."""
    return int(sum(2 * (nj * nk + ni * nk + ni * nj) for ni, nj, nk in dims_list))


# ------------------------------------------------------------------------------ cases
def sod(n=1000, limiter="minmod", riemann="HLLC", cfl=0.5, t_end=0.2):
    """Sod shock tube: x in [0,1], diaphragm at 0.5, (rho,u,p) = (1,0,1) | (0.125,0,0.1);
    faces 1-4 type 400, faces 5-6 type 0 (test/1D/Sod79/INPUT/bc_x*.txt)."""
    dims = (n, 1, 1)
    nodes = cartesian_nodes(dims, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    xc = cell_centres(nodes)[..., 0]
    prim = np.zeros(dims[::-1] + (5,))
    left = xc < 0.5
    prim[..., 0] = np.where(left, 1.0, 0.125)
    prim[..., 4] = np.where(left, 1.0, 0.1)
    blk = Block(dims, nodes, padded_state(dims, prim))
    bc = bc_records([{"dims": dims, "faces": [400, 400, 400, 400, 0, 0]}])
    cfg = abi.make_cfg(GAMMA, RGAS_NONDIM, cfl, limiter, riemann, 3, True)
    return Case("sod", cfg, [blk], bc, t_end=t_end)


def vortex_exact(xc, t, gamma=GAMMA, eps=5.0, x0=5.0, y0=5.0, uinf=(1.0, 1.0, 0.0), L=10.0):
    """Isentropic vortex advected with the free stream (rho,u,v,w,p)=(1,1,1,0,1) over a
    periodic box of side L: returns primitive (..., 5) at points xc (..., 3) and time t."""
    x = xc[..., 0] - uinf[0] * t
    y = xc[..., 1] - uinf[1] * t
    dx = (x - x0 + 0.5 * L) % L - 0.5 * L
    dy = (y - y0 + 0.5 * L) % L - 0.5 * L
    r2 = dx * dx + dy * dy
    f = eps / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    u = uinf[0] - f * dy
    v = uinf[1] + f * dx
    dT = -(gamma - 1.0) * eps * eps / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    Tt = 1.0 + dT
    rho = Tt ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    out = np.empty(xc.shape[:-1] + (5,))
    out[..., 0] = rho
    out[..., 1] = u
    out[..., 2] = v
    out[..., 3] = uinf[2]
    out[..., 4] = p
    return out


def periodic_faces(b):
    return [{"type": 201, "bs": b, "fs": 2}, {"type": 201, "bs": b, "fs": 1},
            {"type": 201, "bs": b, "fs": 4}, {"type": 201, "bs": b, "fs": 3},
            {"type": 201, "bs": b, "fs": 6}, {"type": 201, "bs": b, "fs": 5}]


def vortex(n=256, riemann="HLLC", limiter="vanleer", cfl=0.8, nsteps=200, L=10.0,
           precision=abi.PREC_FP64, dims=None):
    """3-D isentropic vortex on one periodic n^3 block over [0,L]^3, strength 5 at (5,5),
    uniform in z, free stream (1,1,1,0,1); every face type 201 self-periodic."""
    dims = tuple(dims) if dims is not None else (n, n, n)
    nodes = cartesian_nodes(dims, (0.0, 0.0, 0.0), (L, L, L))
    prim = vortex_exact(cell_centres(nodes), 0.0, L=L)
    blk = Block(dims, nodes, padded_state(dims, prim))
    bc = bc_records([{"dims": dims, "faces": periodic_faces(1)}])
    cfg = abi.make_cfg(GAMMA, RGAS_NONDIM, cfl, limiter, riemann, 3, True, precision=precision)
    return Case("vortex", cfg, [blk], bc, nsteps=nsteps, meta={"L": L})


def multiblock_chain(nblocks=64, seed=1234, lo=16, hi=128, h=None, mach=0.8,
                     amp=0.05, riemann="HLLC", limiter="vanleer", cfl=0.5,
                     precision=abi.PREC_FP32_FLUX, max_cells=None):
    """Seeded Cartesian blocks with dims ~ U{lo..hi}, joined in a chain/tree by aligned
    1-to-1 connections (type 101).  Each block is attached to a random earlier block on a
    free face; the two touching faces are forced to equal dims so the connection is
    1-to-1.  Blocks are laid out in space with a common cell size h so connected faces
    coincide.  Free faces are type 400.  State: Mach `mach` free stream along x plus seeded
    smooth Gaussian density/pressure bumps of relative amplitude `amp`."""
    rng = np.random.default_rng(seed)
    dims = [list(rng.integers(lo, hi + 1, size=3)) for _ in range(nblocks)]
    if max_cells is not None:
        while sum(int(np.prod(d)) for d in dims) > max_cells:
            big = max(range(nblocks), key=lambda q: np.prod(dims[q]))
            ax = int(np.argmax(dims[big]))
            dims[big][ax] = max(lo, dims[big][ax] // 2)
    h = h if h is not None else 1.0 / 64.0
    origin = [np.zeros(3) for _ in range(nblocks)]
    faces = [[400] * 6 for _ in range(nblocks)]
    occupied = []          # axis-aligned boxes already placed

    def box(q):
        return origin[q], origin[q] + h * np.array(dims[q], dtype=float)

    def overlaps(q):
        a0, a1 = box(q)
        for r in occupied:
            b0, b1 = box(r)
            if np.all(a0 < b1 - 1e-12) and np.all(b0 < a1 - 1e-12):
                return True
        return False

    occupied.append(0)
    for q in range(1, nblocks):
        placed = False
        for _attempt in range(200):
            parent = int(rng.integers(0, q))
            free = [f for f in range(6) if faces[parent][f] == 400]
            if not free:
                continue
            fp = int(rng.choice(free))          # 0-based face of the parent
            ax = fp // 2
            saved = list(dims[q])
            for t in range(3):                  # match the touching faces
                if t != ax:
                    dims[q][t] = dims[parent][t]
            o = origin[parent].copy()
            if fp % 2 == 1:                     # parent max face -> child on the + side
                o[ax] = origin[parent][ax] + h * dims[parent][ax]
            else:
                o[ax] = origin[parent][ax] - h * dims[q][ax]
            origin[q] = o
            if overlaps(q):
                dims[q] = saved
                continue
            fc = fp + 1 if fp % 2 == 0 else fp - 1     # child face touching the parent
            faces[parent][fp] = {"type": 101, "bs": q + 1, "fs": fc + 1}
            faces[q][fc] = {"type": 101, "bs": parent + 1, "fs": fp + 1}
            occupied.append(q)
            placed = True
            break
        if not placed:
            raise RuntimeError("could not place block %d" % q)
    blocks = []
    a_inf = 1.0
    nb_bumps = 8
    all_lo = np.min([origin[q] for q in range(nblocks)], axis=0)
    all_hi = np.max([origin[q] + h * np.array(dims[q]) for q in range(nblocks)], axis=0)
    centres = all_lo + (all_hi - all_lo) * rng.random((nb_bumps, 3))
    widths = (0.05 + 0.1 * rng.random(nb_bumps)) * np.max(all_hi - all_lo)
    for q in range(nblocks):
        d = tuple(int(v) for v in dims[q])
        nodes = cartesian_nodes(d, origin[q], origin[q] + h * np.array(d, dtype=float))
        xc = cell_centres(nodes)
        bump = np.zeros(xc.shape[:-1])
        for c, w in zip(centres, widths):
            bump += np.exp(-np.sum((xc - c) ** 2, axis=-1) / (w * w))
        prim = np.zeros(xc.shape[:-1] + (5,))
        prim[..., 0] = 1.0 + amp * bump
        prim[..., 1] = mach * a_inf
        prim[..., 4] = (1.0 / GAMMA) * (1.0 + amp * bump)
        blocks.append(Block(d, nodes, padded_state(d, prim)))
    bc = bc_records([{"dims": b.dims, "faces": faces[q]} for q, b in enumerate(blocks)])
    cfg = abi.make_cfg(GAMMA, RGAS_NONDIM, cfl, limiter, riemann, 3, True, precision=precision)
    return Case("multiblock", cfg, blocks, bc, meta={"seed": seed, "faces": faces,
                                                       "origin": origin, "h": h})


def split_block_case(case, axis=0, parts=2):
    """Split the single block of `case` into `parts` blocks along `axis` joined by type-101
    connections; periodic (201) faces of the split axis become 201 connections to the
    opposite end block.  Used to check decomposition invariance of the residual."""
    assert len(case.blocks) == 1
    B = case.blocks[0]
    n = B.dims[axis]
    cuts = np.linspace(0, n, parts + 1).astype(int)
    f_lo, f_hi = 2 * axis + 1, 2 * axis + 2
    old_faces = [None] * 6
    for r in case.bc:
        if old_faces[r["f"] - 1] is None:
            if r["type"] in (101, 201):
                old_faces[r["f"] - 1] = {"type": int(r["type"]), "bs": 1, "fs": int(r["fs"])}
            elif r["type"] == 405:
                old_faces[r["f"] - 1] = {"type": 405, **{k: float(r[k]) for k in
                                                         ("mach", "T", "p", "alpha", "beta", "rel_fac")}}
            else:
                old_faces[r["f"] - 1] = int(r["type"])
    blocks, specs = [], []
    g = abi.GC
    for q in range(parts):
        a, b = cuts[q], cuts[q + 1]
        dims = list(B.dims)
        dims[axis] = b - a
        dims = tuple(dims)
        sl = [slice(None)] * 3
        sl[2 - axis] = slice(a, b + 1)
        nodes = np.ascontiguousarray(B.nodes[tuple(sl)])
        psl = [slice(None)] * 3
        psl[2 - axis] = slice(a, b)
        prim = B.P[g:-g, g:-g, g:-g][tuple(psl)]
        faces = []
        for f in range(1, 7):
            spec = old_faces[f - 1]
            if isinstance(spec, dict) and spec["type"] in (101, 201):
                spec = dict(spec)
                if spec["bs"] == 1 and f not in (f_lo, f_hi):
                    spec["bs"] = q + 1          # self-periodic in another direction
            faces.append(spec)
        if q > 0:
            faces[f_lo - 1] = {"type": 101, "bs": q, "fs": f_hi}
        elif isinstance(old_faces[f_lo - 1], dict) and old_faces[f_lo - 1]["type"] == 201:
            faces[f_lo - 1] = {"type": 201, "bs": parts, "fs": f_hi}
        if q < parts - 1:
            faces[f_hi - 1] = {"type": 101, "bs": q + 2, "fs": f_lo}
        elif isinstance(old_faces[f_hi - 1], dict) and old_faces[f_hi - 1]["type"] == 201:
            faces[f_hi - 1] = {"type": 201, "bs": 1, "fs": f_lo}
        blocks.append(Block(dims, nodes, padded_state(dims, prim)))
        specs.append({"dims": dims, "faces": faces})
    bc = bc_records(specs)
    return Case(case.name + "_split", case.cfg, blocks, bc, t_end=case.t_end, nsteps=case.nsteps,
                meta=dict(case.meta))


def bump_control_shape(xs, cx, width=0.03):
    """Wall displacement of one bump control point of unit height at centre cx."""
    return np.exp(-((xs - cx) / width) ** 2) * ((xs >= 0) & (xs <= 1))


def bump_channel(ni=512, nj=256, height=0.04, mach=2.0, seed=7, ncontrol=64, limiter="vanalbada",
                 riemann="HLLC", cfl=0.5, heights=None, alpha=0.0):
    """2-D supersonic channel (nk=1) over x in [-1.5, 2.5], y in [0, 2], with a circular-arc
    bump on 0<=x<=1 of height `height` (x bump length 1), body-fitted grid.  The wall shape
    is the arc plus a smooth perturbation parametrised by `ncontrol` control-point heights
    (Gaussian bumps, seeded centres); faces: i-min 405 (Mach 2 static p=1/gamma, T=1, flow
    angle alpha from the x axis),
    i-max 400, j-min / j-max 300, k faces 0; local time stepping (steady)."""
    rng = np.random.default_rng(seed)
    dims = (ni, nj, 1)
    xs = np.linspace(-1.5, 2.5, ni + 1)
    ctrl_x = np.sort(rng.random(ncontrol))
    if heights is None:
        heights = np.zeros(ncontrol)
    r = (0.25 + height * height) / (2.0 * height)
    yb = np.where((xs >= 0) & (xs <= 1), np.sqrt(np.maximum(r * r - (xs - 0.5) ** 2, 0.0)) - (r - height), 0.0)
    yb = np.maximum(yb, 0.0)
    for cx, hh in zip(ctrl_x, heights):
        yb = yb + hh * bump_control_shape(xs, cx)
    H = 2.0
    eta = np.linspace(0.0, 1.0, nj + 1)
    X = np.repeat(xs[None, :], nj + 1, axis=0)
    Y = yb[None, :] + eta[:, None] * (H - yb[None, :])
    nodes = np.zeros((2, nj + 1, ni + 1, 3))
    for kk in range(2):
        nodes[kk, :, :, 0] = X
        nodes[kk, :, :, 1] = Y
        nodes[kk, :, :, 2] = float(kk) * 0.1
    a_in = 1.0
    prim = np.zeros((1, nj, ni, 5))
    prim[..., 0] = 1.0
    prim[..., 1] = mach * a_in
    prim[..., 4] = 1.0 / GAMMA
    blk = Block(dims, np.ascontiguousarray(nodes), padded_state(dims, prim))
    inflow = {"type": 405, "mach": mach, "T": 1.0, "p": 1.0 / GAMMA, "alpha": alpha, "beta": 0.0,
              "rel_fac": 1.0}
    bc = bc_records([{"dims": dims, "faces": [inflow, 400, 300, 300, 0, 0]}])
    cfg = abi.make_cfg(GAMMA, RGAS_NONDIM, cfl, limiter, riemann, 3, False)
    return Case("bump", cfg, [blk], bc, meta={"ctrl_x": ctrl_x, "heights": np.asarray(heights, dtype=float),
                                              "xs": xs, "eta": eta, "H": H, "width": 0.03})


def taylor_green(n=192, nbox=8, mach=0.1, seed=11, limiter="vanleer", riemann="AUSM+", cfl=0.5,
                 precision=abi.PREC_FP32_FLUX, viscous=True, pr=0.71, vnn=0.3):
    """Batch of `nbox` periodic n^3 Taylor-Green boxes over [0, 2 pi]^3 at Mach `mach`
    (rho0 = 1, U0 = 1, L = 1, p0 = 1/(gamma M^2), T0 = p0/(rho0 R) = 1/M^2 with R = 1/gamma),
    each box its own self-periodic block with Re drawn from {400, 800, 1600, 3200} by `seed`:
    Sutherland viscosity mu(T0) = 1/Re with S = 110.4/273.15 T0 (air), Pr = `pr`."""
    rng = np.random.default_rng(seed)
    res = rng.choice([400, 800, 1600, 3200], size=nbox)
    blocks, specs, visc = [], [], []
    L = 2.0 * np.pi
    p0 = 1.0 / (GAMMA * mach * mach)
    T0 = p0 / RGAS_NONDIM
    for q in range(nbox):
        dims = (n, n, n)
        nodes = cartesian_nodes(dims, (0.0, 0.0, 0.0), (L, L, L))
        xc = cell_centres(nodes)
        x, y, z = xc[..., 0], xc[..., 1], xc[..., 2]
        prim = np.zeros(xc.shape[:-1] + (5,))
        prim[..., 0] = 1.0
        prim[..., 1] = np.sin(x) * np.cos(y) * np.cos(z)
        prim[..., 2] = -np.cos(x) * np.sin(y) * np.cos(z)
        prim[..., 4] = p0 + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
        blocks.append(Block(dims, nodes, padded_state(dims, prim)))
        specs.append({"dims": dims, "faces": periodic_faces(q + 1)})
        visc.append((q + 1, [1.0 / float(res[q]), T0, 110.4 / 273.15 * T0, pr, vnn]))
    bc = bc_records(specs)
    cfg = abi.make_cfg(GAMMA, RGAS_NONDIM, cfl, limiter, riemann, 3, True, precision=precision)
    meta = {"Re": res, "T0": T0}
    if viscous:
        meta["viscous"] = visc
    return Case("tgv", cfg, blocks, bc, meta=meta)


def kinetic_energy(states):
    """Mean kinetic energy rho |u|^2 / 2 over the cells of each block (uniform cells;
    primitive blk%P-layout states)."""
    out = []
    for P in states:
        Pi = P[2:-2, 2:-2, 2:-2, :]
        out.append(float(np.mean(0.5 * Pi[..., 0] * (Pi[..., 1] ** 2 + Pi[..., 2] ** 2 + Pi[..., 3] ** 2))))
    return out
