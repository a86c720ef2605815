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
    init_p: Optional[List[np.ndarray]] = None  # per block (kdim, jdim, idim) interior p (None: freestream)

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
        """Every block's x, y, z in order (cached: a serving batch rebuilt
        around a request asks for it again)."""
        key = (id(self.xyz), len(self.xyz))
        cached = getattr(self, "_xyz_cache", None)
        if cached is None or cached[0] != key:
            cached = (key, np.concatenate([a.reshape(-1) for a in self.xyz]).astype(np.float64))
            object.__setattr__(self, "_xyz_cache", cached)
        return cached[1]

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
        if self.init_p is not None:
            prim[4, NG:-NG, NG:-NG, NG:-NG] = self.init_p[b]
        return prim

    # ------------------------------------------------------------------ writers
    def block_pids(self, nprocs: int = 1, tries: int = 400) -> List[int]:
        """1-based owner rank of every block for a run of the reference on
        ``nprocs`` MPI ranks.

        The reference tags an interface-residual message with its source
        block (HaloExchanger.cpp:274,303), so two messages from one block to
        two blocks of the same other rank can meet the wrong receives (and
        abort with a truncation error where their sizes differ).  A valid
        placement gives every block at most one interface partner on each
        other rank.  Blocks are placed largest first onto the least-loaded
        rank that keeps the placement valid, with seeded random
        perturbations of both orders over ``tries`` attempts; the most
        balanced valid placement is kept.  Raises ValueError when none is
        found."""
        import random
        nprocs = max(1, int(nprocs))
        nb = self.nblocks
        if nprocs == 1:
            return [1] * nb
        if nprocs > nb:
            raise ValueError("%d ranks for %d blocks" % (nprocs, nb))
        nbrs = [set() for _ in range(nb)]
        for r in self.regions:
            if r.bctype < 0 and r.nbt - 1 != r.block:
                nbrs[r.block].add(r.nbt - 1)
        sizes = [int(v) for v in np.prod(self.dims.astype(np.int64), axis=1)]
        mean = sum(sizes) / nprocs
        bound = max(1.0, max(sizes) / mean)

        def clash(rank, s):
            seen = set()
            for x in nbrs[s]:
                if rank[x] >= 0 and rank[x] != rank[s]:
                    if rank[x] in seen:
                        return True
                    seen.add(rank[x])
            return False

        rng = random.Random(12345)
        best = None
        for t in range(tries):
            jit = 0.0 if t == 0 else 0.3
            rank, load = [-1] * nb, [0] * nprocs
            order = sorted(range(nb), key=lambda b: (-sizes[b] * (1.0 + jit * rng.random()), b))
            for b in order:
                for r in sorted(range(nprocs), key=lambda r: (load[r] + jit * 0.5 * mean * rng.random(), r)):
                    rank[b] = r
                    if not any(rank[s] >= 0 and clash(rank, s) for s in [b] + sorted(nbrs[b])):
                        load[r] += sizes[b]
                        break
                    rank[b] = -1
                if rank[b] < 0:
                    break
            if min(rank) < 0 or any(clash(rank, s) for s in range(nb)):
                continue
            imb = max(load) / mean
            if best is None or imb < best[0]:
                best = (imb, rank)
            if imb <= bound * 1.02:
                break
        if best is None:
            raise ValueError("no placement of %d blocks on %d ranks without a repeated message tag" % (nb, nprocs))
        return [r + 1 for r in best[1]]

    def write(self, casedir: str, nprocs: int = 1, ascii_grid: bool = False) -> None:
        """The case's files; ``nprocs`` > 1 spreads the blocks over that many
        pids in the BC file (block_pids), which the reference's BCReader
        keeps when the run has as many ranks (BCReader.cpp:51-54,80-82);
        ``ascii_grid`` writes the grid in the ASCII form."""
        os.makedirs(casedir, exist_ok=True)
        (write_plot3d_ascii if ascii_grid else write_plot3d)(os.path.join(casedir, "grid.x"), self)
        write_bc(os.path.join(casedir, "grid.inp"), self, self.block_pids(nprocs), max(1, int(nprocs)))
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


def write_plot3d_ascii(path: str, case: Case) -> None:
    """The ASCII grid form the reference falls back to
    (Plot3DReader.cpp:162-208): block count, the sizes, then every block's
    x, y and z, with 17 significant digits (a round trip to the same
    doubles)."""
    with open(path, "w") as f:
        f.write("%d\n" % case.nblocks)
        for b in range(case.nblocks):
            f.write("%d %d %d\n" % tuple(int(v) for v in case.dims[b]))
        for b in range(case.nblocks):
            for c in range(3):
                vals = case.xyz[b][c].reshape(-1)
                for s in range(0, vals.size, 4):
                    f.write(" ".join("%.17g" % v for v in vals[s:s + 4]) + "\n")


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


def write_bc(path: str, case: Case, pids: Optional[Sequence[int]] = None, nprocs: int = 1) -> None:
    """The 1976-marker BC file (BCReader.cpp:41-176): ``nprocs`` processes
    in the file and each block's 1-based pid (all on pid 1 by default)."""
    pids = list(pids) if pids is not None else [1] * case.nblocks
    by_block: Dict[int, List[Region]] = {b: [] for b in range(case.nblocks)}
    for r in case.regions:
        by_block[r.block].append(r)
    lines = ["    1976", f"{int(nprocs):8d}", f"{case.nblocks:8d}"]
    for b in range(case.nblocks):
        i, j, k = (int(v) for v in case.dims[b])
        lines.append(f"{int(pids[b]):8d}")
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
  height = {d(p.get('height', -1.0))}
  tref = {d(p['tref'])}
  pref = {d(p.get('pref', -1.0))}
  rref = {d(p.get('rref', 1.225))}
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


def _flag_pair(rng, own, other, axes):
    """Seeded re-expression of one interface window pair that keeps its
    (identity) orientation, in forms the reference's BC files use
    (testcases/NS/grid_64proc_out.inp:12-15): a tangential axis written with
    negative indices on both sides (-1 .. -n), or run from its high end to
    its low end on both sides (n .. 1).  Never both on one axis: the
    reference's fill_bc_flag loops over the signed window bounds
    (FlowField.cpp:164-196), and -n .. -1 writes outside its arrays.
    own/other: [s_st, s_ed] of the two windows (lists of 3)."""
    for ax in axes:
        u = rng.random()
        if u < 0.45:  # negative flag on both sides
            for w in (own, other):
                w[0][ax], w[1][ax] = -w[0][ax], -w[1][ax]
        elif u < 0.7:  # reversed on both sides
            for w in (own, other):
                w[0][ax], w[1][ax] = w[1][ax], w[0][ax]


def _halo_count(win, dims):
    """Points of one side of the reference's halo message for an interface
    window (get_memory_window, HaloExchanger.cpp:17-64): the tangential
    ranges grow by one node at each end that is not a block edge."""
    n = 1
    for m in range(3):
        a, b = abs(win[0][m]), abs(win[1][m])
        if a == b:
            continue
        lo, hi = min(a, b), max(a, b)
        n *= (hi - lo + 1) + (lo > 1) + (hi < dims[m])
    return n


def column_case(seed: int = 3, ncol=(4, 4), per_column: int = 4, lo: int = 17, hi: int = 129,
                target: Optional[float] = 8.0e6, tol: float = 0.02, tries: int = 100000, *, amp: float = 0.05,
                moo: float = 0.5, reynolds: float = 5.0e3, cfl: float = 5.0, noise: float = 1.0e-3,
                nsteps: int = 1, flags: bool = True, name: str = "columns") -> Case:
    """Blocks on a seeded adjacency graph with partial-face windows.

    ncol[0] x ncol[1] columns in (j, k); column (b, c) has j size sy[b] and
    k size sz[c] (each log-uniform in [lo, hi]) and is split along i into
    ``per_column`` blocks.  The block edges of neighbouring columns
    interleave strictly: a global sequence of 2 (per_column + 1) positions
    is laid out with gaps log-uniform in [lo/2, hi/2]; columns with b + c
    even put their edges near the even positions, the others near the odd
    ones, each edge moved by its own seeded jitter.  So block i sizes fall
    near [lo, hi], columns start and end at different i, and every block
    face meets the neighbouring column's blocks in partial windows, none
    containing another.  That matters to the reference's halo, which grows
    each side's window by one node at every end that is not a block edge
    and unpacks without a map (HaloExchanger.cpp:17-64,161-172): with
    interleaved edges each window pair has one such end per side, the two
    messages have the same length, and the variables stay aligned.  (A
    window that contains its partner's, or edges that coincide, give
    messages of different lengths and the reference then mixes variables;
    the generator checks that no pair does.)  The seed is advanced until
    the node count is within ``tol`` of ``target`` (None: the first seed).
    Blocks share their face node planes.

    Regions of each block, face by face (imin, imax, jmin, jmax, kmin,
    kmax) and along each face by position:
      * i faces: full-face 1:1 windows to the column's next block; the
        column's two ends are far-field (4);
      * j and k faces toward a neighbouring column: a 1:1 window for every
        block of that column overlapping this one by at least two nodes,
        and the parts of the face beyond the neighbouring column's ends a
        physical region of type 4 or 6 (seeded), at least two nodes wide
        (it shares its first or last node line with the window beside it);
      * the global j=1 face is a viscous wall (2), other outer faces
        far-field (4).
    With ``flags`` each interface pair is re-expressed, seeded, with
    negative tangential indices or reversed ranges on both sides
    (_flag_pair); the orientation stays the identity.  Nodes are Cartesian
    plus a seeded sinusoidal perturbation of ``amp`` x spacing."""
    nj_col, nk_col = ncol
    npos = 2 * (per_column + 1)
    for s in range(seed, seed + tries):
        rng = np.random.default_rng(s)
        logu = lambda a, b, n: np.exp(rng.uniform(np.log(a), np.log(b), size=n))  # noqa: E731
        sy = np.rint(logu(lo, hi, nj_col)).astype(int)
        sz = np.rint(logu(lo, hi, nk_col)).astype(int)
        gaps = np.rint(logu(lo / 2.0, hi / 2.0, npos - 1)).astype(int)
        gaps = np.maximum(gaps, 8)
        G = np.concatenate([[0], np.cumsum(gaps)])
        # jitter of each position: strictly inside half the smaller gap beside it,
        # keeping edges of neighbouring columns at least 3 nodes apart
        room = np.minimum(np.concatenate([[gaps[0]], gaps]), np.concatenate([gaps, [gaps[-1]]]))
        jmax = np.maximum((room - 3) // 2, 0)
        edges = np.zeros((nj_col, nk_col, per_column + 1), dtype=int)
        for b in range(nj_col):
            for c in range(nk_col):
                par = (b + c) % 2
                for k in range(per_column + 1):
                    q = 2 * k + par
                    edges[b, c, k] = G[q] + rng.integers(-jmax[q], jmax[q] + 1)
        edges -= edges.min()
        sx = edges[:, :, 1:] - edges[:, :, :-1] + 1
        tot = sum(int(sx[b, c].sum() - (per_column - 1)) * int(sy[b]) * int(sz[c])
                  for b in range(nj_col) for c in range(nk_col))
        if target is None or abs(tot - target) <= tol * target:
            used = s
            break
    else:
        raise ValueError("no seed in [%d, %d) gives %.3g nodes" % (seed, seed + tries, target))
    rng = np.random.default_rng(used + 1000003)
    oy = np.concatenate([[0], np.cumsum(sy - 1)])
    oz = np.concatenate([[0], np.cumsum(sz - 1)])
    NX = int(edges.max()) + 1
    NY, NZ = int(oy[-1]) + 1, int(oz[-1]) + 1
    h = np.array([2.0 / (NX - 1), 1.0 / (NY - 1), 1.0 / (NZ - 1)])
    phases = rng.uniform(0.0, 2 * np.pi, size=(3, 3))
    waves = rng.integers(1, 4, size=(3, 3))

    def coords(I, J, K):
        out = []
        base = [I * h[0], J * h[1], K * h[2]]
        for m in range(3):
            arg = (2 * np.pi * (waves[m, 0] * I / (NX - 1) + waves[m, 1] * J / (NY - 1)
                                + waves[m, 2] * K / (NZ - 1)) + phases[m, 0])
            out.append(base[m] + amp * h[m] * np.sin(arg) * np.cos(phases[m, 1] + 0.5 * arg))
        return out

    blocks = [(a, b, c) for c in range(nk_col) for b in range(nj_col) for a in range(per_column)]
    index = {abc: n for n, abc in enumerate(blocks)}
    dims = np.array([[sx[b, c, a], sy[b], sz[c]] for (a, b, c) in blocks], dtype=np.int32)
    xyz, rhos = [], []
    for n, (a, b, c) in enumerate(blocks):
        ni, nj, nk = (int(v) for v in dims[n])
        K, J, I = np.meshgrid(oz[c] + np.arange(nk), oy[b] + np.arange(nj), edges[b, c, a] + np.arange(ni),
                              indexing="ij")
        xyz.append(np.stack(coords(I.astype(float), J.astype(float), K.astype(float))))
        rhos.append(1.0 + noise * rng.uniform(-1.0, 1.0, size=(nk, nj, ni)))

    FACE = {"imin": 0, "imax": 1, "jmin": 2, "jmax": 3, "kmin": 4, "kmax": 5}
    entries: List[List[tuple]] = [[] for _ in blocks]  # (face, position, region)

    def window(n, face, rng_t=None):
        d = [int(v) for v in dims[n]]
        ax = {"i": 0, "j": 1, "k": 2}[face[0]]
        st, ed = [1, 1, 1], list(d)
        st[ax] = ed[ax] = 1 if face.endswith("min") else d[ax]
        if rng_t is not None:
            st[0], ed[0] = rng_t
        return [st, ed], ax

    pairs = []

    def add_pair(n, face_n, m, face_m, rng_n=None, rng_m=None):
        wn, ax = window(n, face_n, rng_n)
        wm, _ = window(m, face_m, rng_m)
        if _halo_count(wn, dims[n]) != _halo_count(wm, dims[m]):
            raise AssertionError("halo messages of different lengths between blocks %d and %d" % (n + 1, m + 1))
        if flags:
            _flag_pair(rng, wn, wm, [t for t in range(3) if t != ax])
        rn = Region(n, -1, tuple(wn[0]), tuple(wn[1]), tuple(wm[0]), tuple(wm[1]), nbt=m + 1)
        rm = Region(m, -1, tuple(wm[0]), tuple(wm[1]), tuple(wn[0]), tuple(wn[1]), nbt=n + 1)
        entries[n].append((FACE[face_n], rng_n[0] if rng_n else 0, rn))
        entries[m].append((FACE[face_m], rng_m[0] if rng_m else 0, rm))
        pairs.append((rn, rm))

    def add_phys(n, face, kind, rng_t=None):
        w, _ = window(n, face, rng_t)
        entries[n].append((FACE[face], rng_t[0] if rng_t else 0, Region(n, kind, tuple(w[0]), tuple(w[1]))))

    unmatched_kind = lambda: 4 if rng.random() < 0.5 else 6  # noqa: E731
    covered = {}  # (block, face) -> i ranges (local, 1-based) of its windows
    for n, (a, b, c) in enumerate(blocks):
        if a == 0:
            add_phys(n, "imin", 4)
        if a + 1 < per_column:
            add_pair(n, "imax", index[(a + 1, b, c)], "imin")
        else:
            add_phys(n, "imax", 4)
        a0 = int(edges[b, c, a])
        a1 = a0 + int(dims[n][0]) - 1
        for face, other, (b2, c2) in (("jmax", "jmin", (b + 1, c)), ("kmax", "kmin", (b, c + 1))):
            if b2 >= nj_col or c2 >= nk_col:
                continue
            for a2 in range(per_column):
                m = index[(a2, b2, c2)]
                e0 = int(edges[b2, c2, a2])
                e1 = e0 + int(dims[m][0]) - 1
                o0, o1 = max(a0, e0), min(a1, e1)
                if o1 - o0 >= 1:
                    rn, rm = (o0 - a0 + 1, o1 - a0 + 1), (o0 - e0 + 1, o1 - e0 + 1)
                    add_pair(n, face, m, other, rn, rm)
                    covered.setdefault((n, face), []).append(rn)
                    covered.setdefault((m, other), []).append(rm)
    for n, (a, b, c) in enumerate(blocks):
        ni = int(dims[n][0])
        for face, outer, kind in (("jmin", b == 0, 2), ("jmax", b + 1 == nj_col, 4),
                                  ("kmin", c == 0, 4), ("kmax", c + 1 == nk_col, 4)):
            if outer:
                add_phys(n, face, kind)
                continue
            got = sorted(covered.get((n, face), []))
            if not got:
                add_phys(n, face, unmatched_kind())
                continue
            if got[0][0] > 1:
                add_phys(n, face, unmatched_kind(), (1, max(2, got[0][0])))
            for (l0, l1), (r0, r1) in zip(got, got[1:]):
                if r0 > l1:
                    raise AssertionError("gap inside a face of block %d" % (n + 1))
            if got[-1][1] < ni:
                add_phys(n, face, unmatched_kind(), (min(got[-1][1], ni - 1), ni))
    regions: List[Region] = []
    for n in range(len(blocks)):
        entries[n].sort(key=lambda e: (e[0], e[1]))
        regions.extend(r for _, _, r in entries[n])
    where, count = {}, {}
    for r in regions:
        count[r.block] = count.get(r.block, 0) + 1
        where[id(r)] = count[r.block]
    for rn, rm in pairs:
        rn.ibcwin = where[id(rm)]
        rm.ibcwin = where[id(rn)]
    inputs = dict(DEFAULT_INPUTS, moo=moo, reynolds=reynolds, cfl=cfl)
    return Case("%s_seed%d" % (name, used), dims, xyz, regions, inputs, nsteps, rhos)


def split_along_i(case: Case, nblocks: int, name: Optional[str] = None) -> Case:
    """One block cut along i into ``nblocks`` blocks that share their cut
    planes (point-matched), joined by 1:1 full-face windows.  The physical
    regions and the block's self-interfaces (periodic windows of a constant
    k or j plane) are split with it; the i-end regions go to the first and
    last blocks.  Each piece keeps at least 9 nodes along i."""
    if case.nblocks != 1:
        raise ValueError("split_along_i takes a single-block case")
    ni, nj, nk = (int(v) for v in case.dims[0])
    nb = max(1, min(int(nblocks), (ni - 1) // 8))
    cuts = [int(round(q)) for q in np.linspace(0, ni - 1, nb + 1)]
    x = case.xyz[0]
    rho = case.init_rho[0] if case.init_rho is not None else None
    pres = case.init_p[0] if case.init_p is not None else None
    dims, xyz, rhos, ps = [], [], [], []
    for b in range(nb):
        i0, i1 = cuts[b], cuts[b + 1]
        dims.append([i1 - i0 + 1, nj, nk])
        xyz.append(np.ascontiguousarray(x[:, :, :, i0:i1 + 1]))
        if rho is not None:
            rhos.append(np.ascontiguousarray(rho[:, :, i0:i1 + 1]))
        if pres is not None:
            ps.append(np.ascontiguousarray(pres[:, :, i0:i1 + 1]))
    per: List[List[Region]] = [[] for _ in range(nb)]
    for r in case.regions:
        st, ed = list(r.s_st), list(r.s_ed)
        if st[0] == ed[0]:  # an i-end face
            b = 0 if st[0] == 1 else nb - 1
            if b == nb - 1:
                st[0] = ed[0] = dims[b][0]
            if r.bctype < 0:
                raise ValueError("split_along_i: interfaces on the i ends are not split")
            per[b].append(Region(b, r.bctype, tuple(st), tuple(ed)))
            continue
        for b in range(nb):
            i0 = cuts[b]
            s2, e2 = list(st), list(ed)
            s2[0], e2[0] = 1, dims[b][0]
            if r.bctype < 0:  # a self-interface of the block: of each piece with itself
                t2, u2 = list(r.t_st), list(r.t_ed)
                t2[0], u2[0] = 1, dims[b][0]
                per[b].append(Region(b, -1, tuple(s2), tuple(e2), tuple(t2), tuple(u2), nbt=b + 1, ibcwin=0))
            else:
                per[b].append(Region(b, r.bctype, tuple(s2), tuple(e2)))
    # the cut planes
    for b in range(nb - 1):
        d0, d1 = dims[b], dims[b + 1]
        ra = Region(b, -1, (d0[0], 1, 1), (d0[0], nj, nk), (1, 1, 1), (1, nj, nk), nbt=b + 2)
        rb = Region(b + 1, -1, (1, 1, 1), (1, nj, nk), (d0[0], 1, 1), (d0[0], nj, nk), nbt=b + 1)
        per[b].append(ra)
        per[b + 1].insert(0, rb)
    # face order within each block: imin, imax, jmin, jmax, kmin, kmax
    def face_key(r: Region, d):
        for ax in range(3):
            if r.s_st[ax] == r.s_ed[ax]:
                return 2 * ax + (0 if r.s_st[ax] == 1 and d[ax] > 1 else 1)
        return 6
    for b in range(nb):
        per[b].sort(key=lambda r: face_key(r, dims[b]))
    # partner windows: interfaces between pieces and each piece's self-interfaces
    for b in range(nb):
        for r in per[b]:
            if r.bctype >= 0:
                continue
            m = r.nbt - 1
            for w, pr in enumerate(per[m]):
                if pr is not r or m == b:
                    if (pr.bctype < 0 and pr.nbt == b + 1 and tuple(pr.s_st) == tuple(r.t_st)
                            and tuple(pr.s_ed) == tuple(r.t_ed)):
                        r.ibcwin = w + 1
                        break
            else:
                raise RuntimeError("split_along_i: unmatched interface")
    regions = [r for regs in per for r in regs]
    return Case(name or "%s_x%d" % (case.name, nb), np.array(dims, dtype=np.int32), xyz, regions,
                dict(case.inputs), case.nsteps, rhos or None, ps or None)


def bump_request(seed: int, nodes: float, nblocks: int, moo: float, reynolds: float, bump: float,
                 upper_wall: bool, nsteps: int = 1, cfl: float = 5.0) -> Case:
    """A steady channel-bump request: bump_channel_case of about ``nodes``
    interior nodes cut along i into ``nblocks`` point-matched blocks."""
    c = bump_channel_case(seed, nodes, moo, reynolds, bump, upper_wall=upper_wall, nsteps=nsteps, cfl=cfl)
    return split_along_i(c, nblocks) if nblocks > 1 else c


def capacity_case(seed: int = 5, nsteps: int = 1, nbi: int = 4, nbk: int = 12, block=(257, 129, 21),
                  bump: float = 0.08, moo: float = 0.85, reynolds: float = 1.0e6, cfl: float = 20.0,
                  amp: float = 0.05) -> Case:
    """The capacity workload: nbi x nbk blocks of ``block`` nodes (48 of
    257x129x21, 33.4M nodes) forming a transonic channel over a cosine bump
    (height ``bump`` on the lower wall over the middle of the channel):
    blocks cut along i and k, joined by point-matched 1:1 full-face
    windows; viscous lower wall, far-field inflow, outflow and upper
    boundary, symmetry at the two k ends; nodes perturbed by a seeded
    sinusoid of ``amp`` x spacing (zero on the lower wall and the k ends);
    M 0.85, Re 1e6, CFL 20, seeded 1e-3 density noise."""
    bi, bj, bk = block
    NI, NK = nbi * (bi - 1) + 1, nbk * (bk - 1) + 1
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0.0, 2 * np.pi, size=3)
    Lx, Lz = 4.0, 4.0 * (NK - 1) / (NI - 1)
    g = 1.03 ** np.arange(bj - 1)
    eta = np.concatenate([[0.0], np.cumsum(g)]) / g.sum()

    def coords(I, J, K):
        x = I * (Lx / (NI - 1))
        z = K * (Lz / (NK - 1))
        t = np.clip((x - 1.5) / 1.0, 0.0, 1.0)
        ylow = bump * 0.5 * (1.0 - np.cos(2.0 * np.pi * t))
        e = eta[J.astype(int)]
        y = ylow + e * (1.0 - ylow)
        damp = e * (1.0 - e) * np.sin(np.pi * K / (NK - 1))
        hx, hy = Lx / (NI - 1), 1.0 / (bj - 1)
        x = x + amp * hx * damp * np.sin(2 * np.pi * 3 * I / (NI - 1) + ph[0])
        y = y + amp * hy * damp * np.sin(2 * np.pi * 2 * K / (NK - 1) + ph[1])
        return x, y, z

    blocks = [(a, c) for c in range(nbk) for a in range(nbi)]
    dims = np.array([[bi, bj, bk]] * len(blocks), dtype=np.int32)
    xyz, rhos, per = [], [], []
    for n, (a, c) in enumerate(blocks):
        K, J, I = np.meshgrid(c * (bk - 1) + np.arange(bk), np.arange(bj), a * (bi - 1) + np.arange(bi),
                              indexing="ij")
        xyz.append(np.stack(coords(I.astype(float), J, K.astype(float))))
        rhos.append(1.0 + 1.0e-3 * rng.uniform(-1.0, 1.0, size=(bk, bj, bi)))
    index = {abc: n for n, abc in enumerate(blocks)}
    for n, (a, c) in enumerate(blocks):
        regs = []
        for face, nbr, ax, fix, kind in (("imin", (a - 1, c), 0, 1, 4), ("imax", (a + 1, c), 0, bi, 4),
                                         ("jmin", None, 1, 1, 2), ("jmax", None, 1, bj, 4),
                                         ("kmin", (a, c - 1), 2, 1, 3), ("kmax", (a, c + 1), 2, bk, 3)):
            st, ed = [1, 1, 1], [bi, bj, bk]
            st[ax] = ed[ax] = fix
            if nbr is not None and nbr in index:
                tst, ted = list(st), list(ed)
                tst[ax] = ted[ax] = (dims[0][ax] if fix == 1 else 1)
                regs.append(Region(n, -1, tuple(st), tuple(ed), tuple(tst), tuple(ted), nbt=index[nbr] + 1))
            else:
                regs.append(Region(n, kind, tuple(st), tuple(ed)))
        per.append(regs)
    for n, regs in enumerate(per):
        for r in regs:
            if r.bctype < 0:
                for w, pr in enumerate(per[r.nbt - 1]):
                    if pr.bctype < 0 and pr.nbt == n + 1 and tuple(pr.s_st) == tuple(r.t_st):
                        r.ibcwin = w + 1
                        break
    regions = [r for regs in per for r in regs]
    inputs = dict(DEFAULT_INPUTS, moo=moo, reynolds=reynolds, cfl=cfl)
    return Case("capacity_%dx%dx%dx%d" % (len(blocks), bi, bj, bk), dims, xyz, regions, inputs, nsteps, rhos)
