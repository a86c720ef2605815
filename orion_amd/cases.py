"""Synthetic multi-block cases and the reference's on-disk case formats.

A :class:`Case` is everything one solve needs: interior block sizes, node
coordinates, the boundary/interface regions in file order, the namelist
inputs the hot path reads, and an optional initial primitive state.

Writers produce the files the reference reads, byte-compatible with its
readers:
  * Plot3D multi-block grid as Fortran sequential unformatted records with
    int32 markers (src/mesh/Plot3DReader.cpp:17-156);
  * the ``1976`` multi-block BC/topology file
    (src/bc/BCReader.cpp:31-182);
  * the ``param.dat`` namelist (src/core/Params.cpp:103-275);
  * ``init.bin``: prim[6][P] per block, the initial-state override the
    reference harness applies after ``update_temperature``.

Generators follow the workloads the benchmarks are quoted on:
  * :func:`channel_wall_case` -- one 65x33x9 block, geometric wall
    stretching 1.05, viscous wall at j=1, far-field elsewhere, periodic k
    (self-interface), M=0.5, Re=5e3, CFL 5, seeded Gaussian density bump;
  * :func:`lattice_case` -- structured lattices of blocks with 1:1 full-face
    interfaces, Cartesian spacing plus a seeded sinusoidal node
    perturbation, viscous wall at global j=1, far-field elsewhere:
    ``eight_block_case`` (8 x 129x65x65, M=0.8, Re=1e5, CFL 10, seeded 1e-3
    density noise) and ``uneven_lattice_case`` (4x4x4 lattice, per-axis
    sizes log-uniform in [17,129], about 8M nodes).
"""
from __future__ import annotations

import dataclasses
import os
import struct
from typing import Dict, List, Optional, Sequence

import numpy as np

NG = 3
REGION_INTS = 16  # block, bctype, s_st[3], s_ed[3], t_st[3], t_ed[3], nbt, ibcwin


@dataclasses.dataclass
class Region:
    block: int  # 0-based owner
    bctype: int
    s_st: Sequence[int]
    s_ed: Sequence[int]
    t_st: Sequence[int] = (0, 0, 0)
    t_ed: Sequence[int] = (0, 0, 0)
    nbt: int = -1  # 1-based target block for bctype < 0
    ibcwin: int = 0  # 1-based target window for bctype < 0

    def as_ints(self) -> List[int]:
        return ([self.block, self.bctype] + list(self.s_st) + list(self.s_ed) + list(self.t_st)
                + list(self.t_ed) + [self.nbt, self.ibcwin])


DEFAULT_INPUTS = dict(moo=0.5, reynolds=5.0e3, tref=288.15, twall=0.0, cfl=5.0, attack=0.0,
                      sideslip=0.0, csrv=1, nvis=1, method=1, ntmst=0, gcl=1)


@dataclasses.dataclass
class Case:
    name: str
    dims: np.ndarray  # (nb, 3) int32 interior sizes (idim, jdim, kdim)
    xyz: List[np.ndarray]  # per block (3, kdim, jdim, idim) float64 (i fastest)
    regions: List[Region]
    inputs: Dict[str, float]
    nsteps: int = 1
    init_rho: Optional[List[np.ndarray]] = None  # per block (kdim, jdim, idim) interior rho

    @property
    def nblocks(self) -> int:
        return int(self.dims.shape[0])

    def padded(self, b: int):
        i, j, k = (int(v) for v in self.dims[b])
        return (i + 2 * NG, j + 2 * NG, k + 2 * NG)

    def interior_nodes(self) -> int:
        return int(np.prod(self.dims.astype(np.int64), axis=1).sum())

    def padded_points(self) -> int:
        return int(sum(np.prod(self.padded(b)) for b in range(self.nblocks)))

    def region_table(self) -> np.ndarray:
        return np.asarray([r.as_ints() for r in self.regions], dtype=np.int32).reshape(-1, REGION_INTS)

    def xyz_flat(self) -> np.ndarray:
        return np.concatenate([a.reshape(-1) for a in self.xyz]).astype(np.float64)

    def freestream(self):
        """Freestream primitive state (Inflow.cpp init_inflow, custom atmosphere)."""
        g = 1.40
        moo = float(self.inputs["moo"])
        a = float(self.inputs.get("attack", 0.0)) * 3.14159265358979323846 / 180.0
        s = float(self.inputs.get("sideslip", 0.0)) * 3.14159265358979323846 / 180.0
        import math
        u = math.cos(a) * math.cos(s)
        v = math.sin(a) * math.cos(s)
        w = math.sin(s)
        p = 1.0 / (g * moo * moo)
        return np.array([1.0, u, v, w, p, 1.0])

    def initial_prim(self, b: int) -> np.ndarray:
        """prim[6, nz, ny, nx] before update_temperature: freestream everywhere,
        with the seeded density field on interior nodes."""
        nx, ny, nz = self.padded(b)
        prim = np.empty((6, nz, ny, nx))
        prim[:] = self.freestream()[:, None, None, None]
        if self.init_rho is not None:
            prim[0, NG:-NG, NG:-NG, NG:-NG] = self.init_rho[b]
        return prim

    # ------------------------------------------------------------------ writers
    def write(self, casedir: str) -> None:
        os.makedirs(casedir, exist_ok=True)
        write_plot3d(os.path.join(casedir, "grid.x"), self)
        write_bc(os.path.join(casedir, "grid.inp"), self)
        write_param(os.path.join(casedir, "param.dat"), self)
        with open(os.path.join(casedir, "init.bin"), "wb") as f:
            for b in range(self.nblocks):
                f.write(np.ascontiguousarray(self.initial_prim(b)).tobytes())


def _rec(f, payload: bytes) -> None:
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def write_plot3d(path: str, case: Case) -> None:
    """Multi-block Plot3D, Fortran sequential records (Plot3DReader.cpp:88-156)."""
    with open(path, "wb") as f:
        _rec(f, struct.pack("<i", case.nblocks))
        for b in range(case.nblocks):
            _rec(f, np.asarray(case.dims[b], dtype="<i4").tobytes())
        for b in range(case.nblocks):
            _rec(f, np.ascontiguousarray(case.xyz[b], dtype="<f8").tobytes())


def read_plot3d(path: str):
    """Inverse of :func:`write_plot3d` (double records only)."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0

    def rec():
        nonlocal pos
        (n,) = struct.unpack_from("<i", data, pos)
        payload = data[pos + 4: pos + 4 + n]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        if n != n2:
            raise ValueError("Fortran record length mismatch")
        pos += 8 + n
        return payload

    (nb,) = struct.unpack("<i", rec())
    dims = np.array([np.frombuffer(rec(), dtype="<i4") for _ in range(nb)], dtype=np.int32)
    xyz = []
    for b in range(nb):
        i, j, k = dims[b]
        xyz.append(np.frombuffer(rec(), dtype="<f8").reshape(3, k, j, i).copy())
    return dims, xyz


def write_bc(path: str, case: Case) -> None:
    """The 1976-marker BC file (BCReader.cpp:41-176), all blocks on pid 1."""
    by_block: Dict[int, List[Region]] = {b: [] for b in range(case.nblocks)}
    for r in case.regions:
        by_block[r.block].append(r)
    lines = ["    1976", "       1", f"{case.nblocks:8d}"]
    for b in range(case.nblocks):
        i, j, k = (int(v) for v in case.dims[b])
        lines.append("       1")
        lines.append(f"{i:8d}{j:8d}{k:8d}")
        lines.append(f" blk-{b + 1}")
        lines.append(f"{len(by_block[b]):8d}")
        for r in by_block[b]:
            s = (r.s_st[0], r.s_ed[0], r.s_st[1], r.s_ed[1], r.s_st[2], r.s_ed[2])
            lines.append("".join(f"{v:8d}" for v in s) + f"{r.bctype:8d}")
            if r.bctype < 0:
                t = (r.t_st[0], r.t_ed[0], r.t_st[1], r.t_ed[1], r.t_st[2], r.t_ed[2])
                lines.append("        " + "".join(f"{v:8d}" for v in t) + f"{r.nbt:8d}{r.ibcwin:8d}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_param(path: str, case: Case) -> None:
    """param.dat namelist (Params.cpp:103-275): custom atmosphere (height<0)
    with tref and rref given, so init_inflow keeps the input Reynolds number
    and sets visc = 110.4/tref."""
    p = dict(DEFAULT_INPUTS)
    p.update(case.inputs)

    def d(x):
        return repr(float(x))

    text = f"""$inflow
  moo = {d(p['moo'])}
  attack = {d(p['attack'])}
  sideslip = {d(p['sideslip'])}
  reynolds = {d(p['reynolds'])}
  twall = {d(p['twall'])}
  height = -1.0
  tref = {d(p['tref'])}
  pref = -1.0
  rref = 1.225
  ndim = 3
/
$force
  nwholefield = 1
  lfref = 1.0
  lref = 1.0
/
$filename
  gridname = 'grid.x'
  bcname = 'grid.inp'
  flowname = 'output/'
/
$control
  nstart = 0
  nomax = {int(case.nsteps)}
  method = {int(p['method'])}
/
$step
  ntmst = {int(p['ntmst'])}
  cfl = {d(p['cfl'])}
/
$flowtype
  nvis = {int(p['nvis'])}
/
$technic
  nscheme = 41
  csrv = {int(p['csrv'])}
/
$gcl_cic
  gcl = {int(p['gcl'])}
/
"""
    with open(path, "w") as f:
        f.write(text)


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------

def _face_regions(b: int, dims, kinds: Dict[str, int]) -> List[Region]:
    """Full-face physical regions for faces listed in ``kinds`` in the order
    imin, imax, jmin, jmax, kmin, kmax."""
    i, j, k = dims
    win = {
        "imin": ((1, 1, 1), (1, j, k)), "imax": ((i, 1, 1), (i, j, k)),
        "jmin": ((1, 1, 1), (i, 1, k)), "jmax": ((1, j, 1), (i, j, k)),
        "kmin": ((1, 1, 1), (i, j, 1)), "kmax": ((1, 1, k), (i, j, k)),
    }
    out = []
    for face in ("imin", "imax", "jmin", "jmax", "kmin", "kmax"):
        if face in kinds:
            st, ed = win[face]
            out.append(Region(b, kinds[face], st, ed))
    return out


def channel_wall_case(seed: int = 1, nsteps: int = 200, dims=(65, 33, 9), moo=0.5, reynolds=5.0e3,
                      cfl=5.0) -> Case:
    """One block, geometric wall stretching 1.05 in j, viscous adiabatic wall
    at j=1, far-field at i=1, i=imax and j=jmax, k=1 <-> k=kmax periodic."""
    ni, nj, nk = dims
    dx = 2.0 / (ni - 1)
    h0 = 1.0 * 0.05 / (1.05 ** (nj - 1) - 1.0)
    y = np.concatenate([[0.0], np.cumsum(h0 * 1.05 ** np.arange(nj - 1))])
    x = np.arange(ni) * dx
    z = np.arange(nk) * dx
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    xyz = [np.stack([X, Y, Z])]
    regs = [
        Region(0, 4, (1, 1, 1), (1, nj, nk)),
        Region(0, 4, (ni, 1, 1), (ni, nj, nk)),
        Region(0, 2, (1, 1, 1), (ni, 1, nk)),
        Region(0, 4, (1, nj, 1), (ni, nj, nk)),
        Region(0, -1, (1, 1, 1), (ni, nj, 1), (1, 1, nk), (ni, nj, nk), nbt=1, ibcwin=6),
        Region(0, -1, (1, 1, nk), (ni, nj, nk), (1, 1, 1), (ni, nj, 1), nbt=1, ibcwin=5),
    ]
    rng = np.random.default_rng(seed)
    x0 = np.array([rng.uniform(0.6, 1.4), rng.uniform(0.15, 0.5), z[nk // 2]])
    sig = rng.uniform(0.15, 0.3)
    r2 = (X - x0[0]) ** 2 + (Y - x0[1]) ** 2 + (Z - x0[2]) ** 2
    rho = 1.0 + 1.0e-2 * np.exp(-r2 / sig ** 2)
    inputs = dict(DEFAULT_INPUTS, moo=moo, reynolds=reynolds, cfl=cfl)
    return Case("channel_wall_%dx%dx%d" % tuple(dims), np.array([dims], dtype=np.int32), xyz, regs,
                inputs, nsteps, [rho])


def lattice_case(sx: Sequence[int], sy: Sequence[int], sz: Sequence[int], *, seed: int = 7,
                 amp: float = 0.05, extent=(2.0, 1.0, 1.0), moo=0.8, reynolds=1.0e5, cfl=10.0,
                 noise: float = 1.0e-3, nsteps: int = 1, name: str = "lattice",
                 shared_faces: bool = True) -> Case:
    """Blocks on a structured lattice; block (a,b,c) has interior sizes
    (sx[a], sy[b], sz[c]) and meets its lattice neighbours through 1:1
    full-face windows.  Nodes are Cartesian plus a seeded sinusoidal
    perturbation of ``amp`` x spacing, a function of the global node index.

    shared_faces=True (default): point-matched blocks sharing their face
    node planes.  shared_faces=False: abutting blocks, the first node of a
    block one spacing beyond the last node of its neighbour (the geometry
    in which the reference's exchange_bc, which sends the sender's face node
    as the receiver's first ghost layer, HaloExchanger.cpp:34-50, is
    consistent).  With either geometry the reference's multi-block solution
    grows without bound within ~100-150 steps at any CFL tried (0.5-10),
    while single blocks converge; the oracle reproduces this bit for bit,
    so it is a property of the reference's interface coupling."""
    rng = np.random.default_rng(seed)
    ov = 1 if shared_faces else 0
    NX = sum(sx) - ov * (len(sx) - 1)
    NY = sum(sy) - ov * (len(sy) - 1)
    NZ = sum(sz) - ov * (len(sz) - 1)
    h = np.array([extent[0] / (NX - 1), extent[1] / (NY - 1), extent[2] / (NZ - 1)])
    phases = rng.uniform(0.0, 2 * np.pi, size=(3, 3))
    waves = rng.integers(1, 4, size=(3, 3))

    def coords(I, J, K):
        base = [I * h[0], J * h[1], K * h[2]]
        out = []
        for c in range(3):
            arg = (2 * np.pi * (waves[c, 0] * I / (NX - 1) + waves[c, 1] * J / (NY - 1)
                                + waves[c, 2] * K / (NZ - 1)) + phases[c, 0])
            out.append(base[c] + amp * h[c] * np.sin(arg) * np.cos(phases[c, 1] + 0.5 * arg))
        return out

    ox = np.concatenate([[0], np.cumsum([s - ov for s in sx])])
    oy = np.concatenate([[0], np.cumsum([s - ov for s in sy])])
    oz = np.concatenate([[0], np.cumsum([s - ov for s in sz])])
    blocks = []
    for c in range(len(sz)):
        for b in range(len(sy)):
            for a in range(len(sx)):
                blocks.append((a, b, c))
    index = {abc: n for n, abc in enumerate(blocks)}
    dims = np.array([[sx[a], sy[b], sz[c]] for (a, b, c) in blocks], dtype=np.int32)
    xyz, rhos = [], []
    per_block: List[List[Region]] = []
    for n, (a, b, c) in enumerate(blocks):
        ni, nj, nk = dims[n]
        K, J, I = np.meshgrid(oz[c] + np.arange(nk), oy[b] + np.arange(nj), ox[a] + np.arange(ni),
                              indexing="ij")
        xyz.append(np.stack(coords(I.astype(float), J.astype(float), K.astype(float))))
        rhos.append(1.0 + noise * rng.uniform(-1.0, 1.0, size=(nk, nj, ni)))
        regs = []
        faces = [("imin", (a - 1, b, c), 0, 1), ("imax", (a + 1, b, c), 0, ni),
                 ("jmin", (a, b - 1, c), 1, 1), ("jmax", (a, b + 1, c), 1, nj),
                 ("kmin", (a, b, c - 1), 2, 1), ("kmax", (a, b, c + 1), 2, nk)]
        for face, nbr, ax, fix in faces:
            st = [1, 1, 1]
            ed = [int(ni), int(nj), int(nk)]
            st[ax] = ed[ax] = fix
            if nbr in index:
                m = index[nbr]
                tdims = dims[m]
                tst = [1, 1, 1]
                ted = [int(v) for v in tdims]
                tfix = int(tdims[ax]) if fix == 1 else 1
                tst[ax] = ted[ax] = tfix
                regs.append(Region(n, -1, tuple(st), tuple(ed), tuple(tst), tuple(ted), nbt=m + 1,
                                   ibcwin=0))
            else:
                kind = 2 if face == "jmin" else 4
                regs.append(Region(n, kind, tuple(st), tuple(ed)))
        per_block.append(regs)
    # resolve ibcwin: partner window whose target is this region
    for n, regs in enumerate(per_block):
        for r in regs:
            if r.bctype < 0:
                m = r.nbt - 1
                for w, pr in enumerate(per_block[m]):
                    if pr.bctype < 0 and pr.nbt == n + 1 and tuple(pr.s_st) == tuple(r.t_st) \
                            and tuple(pr.s_ed) == tuple(r.t_ed):
                        r.ibcwin = w + 1
                        break
                else:
                    raise RuntimeError("unmatched interface")
    regions = [r for regs in per_block for r in regs]
    inputs = dict(DEFAULT_INPUTS, moo=moo, reynolds=reynolds, cfl=cfl)
    return Case(name, dims, xyz, regions, inputs, nsteps, rhos)


def eight_block_case(seed: int = 7, nsteps: int = 500, block=(129, 65, 65)) -> Case:
    """8 blocks of 129x65x65 on a 2x2x2 lattice, M=0.8, Re=1e5, CFL 10."""
    bi, bj, bk = block
    return lattice_case([bi, bi], [bj, bj], [bk, bk], seed=seed, nsteps=nsteps,
                        name="eight_block_%dx%dx%d" % tuple(block))


def uneven_sizes(seed: int, n: int = 4, lo: int = 17, hi: int = 129, target: float = 8.0e6,
                 tol: float = 0.02, tries: int = 100000):
    """Per-axis lattice sizes drawn log-uniformly from [lo, hi]; the seed is
    advanced until the total node count is within ``tol`` of ``target``
    (at most ``tries`` seeds: a target far from what [lo, hi] gives raises).
    Returns (sizes_x, sizes_y, sizes_z, seed_used)."""
    for s in range(seed, seed + tries):
        rng = np.random.default_rng(s)
        sz = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=(3, n)))).astype(int)
        tot = float(sz[0].sum() * sz[1].sum() * sz[2].sum())
        if abs(tot - target) <= tol * target:
            return list(sz[0]), list(sz[1]), list(sz[2]), s
    raise ValueError("no seed in [%d, %d) gives %.3g nodes from sizes in [%d, %d]" % (seed, seed + tries, target, lo, hi))


def uneven_lattice_case(seed: int = 11, nsteps: int = 1, target: float = 8.0e6, lo: int = 17,
                        hi: int = 129) -> Case:
    """64 uneven blocks (4x4x4 lattice, sizes log-U[lo, hi], about ``target``
    nodes; by default 17-129 and 8M), same physics as the channel case
    (M=0.5, Re=5e3, CFL 5)."""
    sx, sy, sz, used = uneven_sizes(seed, lo=lo, hi=hi, target=target)
    c = lattice_case(sx, sy, sz, seed=used, moo=0.5, reynolds=5.0e3, cfl=5.0, nsteps=nsteps,
                     name="uneven64_seed%d" % used)
    return c


def scaled_channel(ni: int, nj: int, nk: int, **kw) -> Case:
    return channel_wall_case(dims=(ni, nj, nk), **kw)


def bump_channel_case(seed: int, nodes: float, moo: float, reynolds: float, bump: float, *,
                      upper_wall: bool = False, nsteps: int = 1, cfl: float = 5.0) -> Case:
    """A steady channel-bump request of the serving stream: one block of
    about ``nodes`` interior nodes (aspect 4 : 2 : 1 in i, j, k), x in
    [0, 2], a cosine bump of height ``bump`` on the lower wall over
    x in [0.6, 1.4], wall stretching 1.05 in j, height 1; viscous adiabatic
    lower wall, far-field inflow and outflow, upper wall viscous or
    far-field, periodic in k.  Single blocks, because the reference's
    multi-block solutions do not converge (see lattice_case)."""
    n = max(float(nodes), 8.0 * 7 * 7 * 7)
    nk = max(7, int(round((n / 8.0) ** (1.0 / 3.0))))
    nj = max(9, 2 * nk)
    ni = max(13, int(round(n / (nj * nk))))
    x = np.linspace(0.0, 2.0, ni)
    t = np.clip((x - 0.6) / 0.8, 0.0, 1.0)
    ylow = bump * 0.5 * (1.0 - np.cos(2.0 * np.pi * t))
    g = 1.05 ** np.arange(nj - 1)
    eta = np.concatenate([[0.0], np.cumsum(g)]) / g.sum()  # 0..1, clustered at the wall
    dz = 2.0 / (ni - 1)
    z = np.arange(nk) * dz
    Z, E, X = np.meshgrid(z, eta, x, indexing="ij")
    YL = np.broadcast_to(ylow[None, None, :], X.shape)
    Y = YL + E * (1.0 - YL)
    xyz = [np.stack([X, Y, Z])]
    regs = [
        Region(0, 4, (1, 1, 1), (1, nj, nk)),
        Region(0, 4, (ni, 1, 1), (ni, nj, nk)),
        Region(0, 2, (1, 1, 1), (ni, 1, nk)),
        Region(0, 2 if upper_wall else 4, (1, nj, 1), (ni, nj, nk)),
        Region(0, -1, (1, 1, 1), (ni, nj, 1), (1, 1, nk), (ni, nj, nk), nbt=1, ibcwin=6),
        Region(0, -1, (1, 1, nk), (ni, nj, nk), (1, 1, 1), (ni, nj, 1), nbt=1, ibcwin=5),
    ]
    # seeded Gaussian density bump of 1e-2, as in channel_wall_case
    rng = np.random.default_rng(seed)
    x0 = np.array([rng.uniform(0.4, 1.6), rng.uniform(0.3, 0.7), z[nk // 2]])
    sig = rng.uniform(0.15, 0.3)
    r2 = (X - x0[0]) ** 2 + (Y - x0[1]) ** 2 + (Z - x0[2]) ** 2
    rho = 1.0 + 1.0e-2 * np.exp(-r2 / sig ** 2)
    inputs = dict(DEFAULT_INPUTS, moo=float(moo), reynolds=float(reynolds), cfl=float(cfl))
    return Case("%s_%d_%dx%dx%d" % ("bump" if bump > 0 else "channel", seed, ni, nj, nk),
                np.array([[ni, nj, nk]], dtype=np.int32), xyz, regs,
                inputs, nsteps, [rho])
