"""Small cases for parity tests (the oracle finishes them in well under a second)."""
import numpy as np

from orion_amd import cases
from orion_amd.cases import Case, Region, DEFAULT_INPUTS


def channel_small(nsteps=3):
    """Wall + far-field + periodic self-interface, 33x17x9."""
    return cases.channel_wall_case(seed=3, nsteps=nsteps, dims=(33, 17, 9))


def lattice_small(nsteps=3):
    """2x2x1 lattice: 1:1 interfaces, wall and far-field, curvilinear grid."""
    return cases.lattice_case([13, 11], [9, 12], [10], seed=5, nsteps=nsteps, name="lattice_small")


def lattice_supersonic(nsteps=3):
    return cases.lattice_case([15, 9], [11], [9, 8], seed=9, moo=2.0, nsteps=nsteps, name="lattice_sup")


def bc_variety(nvis=1, nsteps=3):
    """One block exercising freestream (5), extrapolation (6), symmetry (3)
    and the wall (2, inviscid slip when nvis=0), periodic in k."""
    ni, nj, nk = 19, 13, 9
    dx = 1.0 / (ni - 1)
    x = np.arange(ni) * dx
    y = np.arange(nj) * dx * 0.9
    z = np.arange(nk) * dx * 1.1
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    X = X + 0.02 * dx * np.sin(3.0 * Y / dx) * np.cos(Z / dx)
    xyz = [np.stack([X, Y, Z])]
    regs = [
        Region(0, 5, (1, 1, 1), (1, nj, nk)),
        Region(0, 6, (ni, 1, 1), (ni, nj, nk)),
        Region(0, 2, (1, 1, 1), (ni, 1, nk)),
        Region(0, 3, (1, nj, 1), (ni, nj, nk)),
        Region(0, -1, (1, 1, 1), (ni, nj, 1), (1, 1, nk), (ni, nj, nk), nbt=1, ibcwin=6),
        Region(0, -1, (1, 1, nk), (ni, nj, nk), (1, 1, 1), (ni, nj, 1), nbt=1, ibcwin=5),
    ]
    rng = np.random.default_rng(21)
    rho = 1.0 + 1e-2 * rng.uniform(-1, 1, size=(nk, nj, ni))
    inputs = dict(DEFAULT_INPUTS, moo=0.6, reynolds=2.0e3, cfl=3.0, nvis=nvis)
    return Case("bc_variety_nvis%d" % nvis, np.array([[ni, nj, nk]], dtype=np.int32), xyz, regs, inputs, nsteps,
                [rho])


def isothermal_wall(nsteps=3):
    c = cases.channel_wall_case(seed=4, nsteps=nsteps, dims=(21, 15, 8))
    c.inputs = dict(c.inputs, twall=350.0)
    c.name = "isothermal_wall"
    return c


def short_duct(nsteps=3):
    """Viscous walls at both j ends of a 9-point j extent, far-field at the i
    ends, periodic in k: the boundary closures of a line's two ends share
    faces (the right end reads states the left end's stages set)."""
    ni, nj, nk = 15, 9, 7
    c = cases.channel_wall_case(seed=6, nsteps=nsteps, dims=(ni, nj, nk))
    c.regions[3] = cases.Region(0, 2, (1, nj, 1), (ni, nj, nk))
    c.name = "short_duct"
    return c


ALL = {
    "channel_small": channel_small,
    "lattice_small": lattice_small,
    "lattice_supersonic": lattice_supersonic,
    "bc_variety_visc": lambda nsteps=3: bc_variety(1, nsteps),
    "bc_variety_inviscid": lambda nsteps=3: bc_variety(0, nsteps),
    "isothermal_wall": isothermal_wall,
    "short_duct": short_duct,
}
