"""Python mirror of ORION's hot-path operator surface, backed by the gfx950
library through its C ABI (include/orion_hip.h).

Names, argument order and meaning follow the reference's operators:
  solver::HaloExchanger::exchange_bc / average_interface_residuals
      (src/solver/HaloExchanger.hpp:20,25)
  bc::apply_physical_bc                    (src/bc/PhysicalBC.hpp:17)
  solver::StateUpdater::update_flow_states / update_solution
      (src/solver/StateUpdater.hpp:14,21)
  solver::TimeIntegrator::calculate_time_step (src/solver/TimeIntegrator.hpp:10)
  solver::FluxComputer::compute_viscous_rhs   (src/solver/FluxComputer.hpp:15)
  solver::InviscidFluxComputer::compute_inviscid_rhs
      (src/solver/InviscidFluxComputer.hpp:17)
  solver::ResidualMonitor::compute            (src/solver/ResidualMonitor.hpp:20)
Errors raise (OrionHipError, a RuntimeError), as the reference throws
std::runtime_error.

The field set is device resident; ``field()`` copies a block field to the
host in the reference's OrionArray layout (nvar, nz, ny, nx), i fastest.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from types import SimpleNamespace
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import FIELD_IDS, Inputs, RegionDesc, ResidualStats, check

NG = 3


@dataclasses.dataclass
class ResidualStatsPy:
    rms_residual: float
    max_residual: float
    max_loc: tuple
    phydtime: float


def _stats(s: ResidualStats) -> ResidualStatsPy:
    return ResidualStatsPy(s.rms_residual, s.max_residual, tuple(s.max_loc), s.phydtime)


def make_params(case) -> SimpleNamespace:
    """A core::Params-shaped namespace holding the inputs the path reads."""
    inp = dict(case.inputs)
    return SimpleNamespace(
        inflow=SimpleNamespace(moo=inp["moo"], reynolds=inp["reynolds"], tref=inp["tref"], twall=inp["twall"],
                               attack=inp.get("attack", 0.0), sideslip=inp.get("sideslip", 0.0), gama=1.4,
                               prl=0.72),
        step=SimpleNamespace(cfl=inp["cfl"], ntmst=inp.get("ntmst", 0), phydtime=0.0),
        technic=SimpleNamespace(csrv=int(inp.get("csrv", 1))),
        flowtype=SimpleNamespace(nvis=int(inp.get("nvis", 1))),
        control=SimpleNamespace(method=int(inp.get("method", 1)), nomax=case.nsteps),
        gcl_cic=SimpleNamespace(gcl=int(inp.get("gcl", 1))),
        filename=SimpleNamespace(flowname="output/"),  # what Case.write puts in param.dat
    )


def region_descs(cases, offsets=None):
    """The C-ABI region array of one or several cases' regions, in order,
    with block numbers (owner, interface target) shifted by each case's
    block offset."""
    cases = list(cases)
    offsets = offsets if offsets is not None else [0] * len(cases)
    total = sum(len(c.regions) for c in cases)
    regs = (RegionDesc * max(1, total))()
    n = 0
    for case, off in zip(cases, offsets):
        for r in case.regions:
            ints = r.as_ints()
            regs[n].block, regs[n].bctype = ints[0] + off, ints[1]
            for m in range(3):
                regs[n].s_st[m] = ints[2 + m]
                regs[n].s_ed[m] = ints[5 + m]
                regs[n].t_st[m] = ints[8 + m]
                regs[n].t_ed[m] = ints[11 + m]
            regs[n].nbt = ints[14] + off if ints[14] > 0 else ints[14]
            regs[n].ibcwin = ints[15]
            n += 1
    return regs, total


def case_inputs(case) -> Inputs:
    return Inputs(**{k: case.inputs[k] for k, _ in Inputs._fields_})


class FlowFieldSet:
    """Device mirror of preprocess::FlowFieldSet (src/preprocess/FlowField.hpp:73-87)
    for one case: fields, metrics and the compiled boundary topology."""

    ng = NG
    nvar = 5
    nprim = 6

    def __init__(self, case, keep_q: bool = False, device: int = 0, apply_init: bool = True):
        L = _lib.load()
        check(L.orion_hip_backend_init(int(device)), "backend_init")
        self.case = case
        self.bc = case.regions
        dims = np.ascontiguousarray(case.dims, dtype=np.int32).reshape(-1)
        xyz = np.ascontiguousarray(case.xyz_flat(), dtype=np.float64)
        regs, _ = region_descs([case])
        inp = case_inputs(case)
        h = C.c_void_p()
        check(L.orion_hip_ctx_create(case.nblocks, dims.ctypes.data_as(C.POINTER(C.c_int)),
                                     xyz.ctypes.data_as(C.POINTER(C.c_double)), len(case.regions), regs,
                                     C.byref(inp), 1 if keep_q else 0, C.byref(h)), "ctx_create")
        self.ctx = h
        self.local_block_ids = list(range(case.nblocks))
        if apply_init and case.init_rho is not None:
            for b in range(case.nblocks):
                self.set_field(b, "prim", case.initial_prim(b))
            check(L.orion_hip_update_temperature(self.ctx), "update_temperature")

    # ------------------------------------------------------------------ data
    def close(self) -> None:
        if getattr(self, "ctx", None):
            _lib.load().orion_hip_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, b: int, name: str):
        nx, ny, nz = self.case.padded(b)
        return (FIELD_IDS[name][1], nz, ny, nx)

    def field(self, b: int, name: str) -> np.ndarray:
        fid, nv = FIELD_IDS[name]
        out = np.empty(self.shape(b, name))
        check(_lib.load().orion_hip_download(self.ctx, b, fid, out.ctypes.data_as(C.POINTER(C.c_double)), out.size),
              "download")
        return out

    def set_field(self, b: int, name: str, arr: np.ndarray) -> None:
        fid, nv = FIELD_IDS[name]
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape(b, name))
        check(_lib.load().orion_hip_upload(self.ctx, b, fid, a.ctypes.data_as(C.POINTER(C.c_double)), a.size),
              "upload")

    def device_ptr(self, b: int, name: str) -> int:
        fid, _ = FIELD_IDS[name]
        p = C.c_void_p()
        n = C.c_longlong()
        check(_lib.load().orion_hip_field(self.ctx, b, fid, C.byref(p), C.byref(n)), "field")
        return int(p.value)

    def device_bytes(self) -> int:
        return int(_lib.load().orion_hip_device_bytes(self.ctx))

    def interior_nodes(self) -> int:
        return int(_lib.load().orion_hip_interior_nodes(self.ctx))

    def synchronize(self) -> None:
        check(_lib.load().orion_hip_synchronize(self.ctx), "synchronize")

    def set_stream(self, stream_ptr: Optional[int]) -> None:
        check(_lib.load().orion_hip_set_stream(self.ctx, C.c_void_p(stream_ptr or 0)), "set_stream")

    def set_overlap(self, on: bool) -> None:
        """With viscosity, run the inviscid face fluxes on a second stream
        beside the viscous chain (the default) or the whole step on one
        stream (many contexts busy at once); the results are the same."""
        check(_lib.load().orion_hip_set_overlap(self.ctx, int(bool(on))), "set_overlap")

    def set_per_block(self, on: bool) -> None:
        """The plain per-block form (every kernel once per block, in block
        order, no captured step) or the batched form (the default: each
        kernel over all blocks, the step replayed as a graph); the results
        are the same bit for bit."""
        check(_lib.load().orion_hip_set_per_block(self.ctx, int(bool(on))), "set_per_block")

    def set_visc_tiles(self, segment: int, grid: int) -> None:
        """The viscous face-flux stage's tiling: planes per item (fused
        form) or faces per marching item (queue form), and the workgroups of
        its persistent launch (results do not depend on it)."""
        check(_lib.load().orion_hip_set_visc_tiles(self.ctx, int(segment), int(grid)), "set_visc_tiles")

    def set_visc_form(self, fused: bool) -> None:
        """The viscous right-hand side after the node derivatives: fused
        (face fluxes and their difference in one pass; every block line
        needs at least 4 nodes) or the face-flux queue and a difference
        kernel (the default); the results are the same bit for bit."""
        check(_lib.load().orion_hip_set_visc_form(self.ctx, int(bool(fused))), "set_visc_form")

    def visc_form(self) -> bool:
        return bool(_lib.load().orion_hip_get_visc_form(self.ctx))

    def visc_tiles(self):
        seg, grid = C.c_int(), C.c_int()
        check(_lib.load().orion_hip_get_visc_tiles(self.ctx, C.byref(seg), C.byref(grid)), "get_visc_tiles")
        return seg.value, grid.value

    def tune(self, steps: int = 3) -> dict:
        """The launch-shape autotuner (orion_hip_tune): sweep schedule, then
        viscous segment, then its grid, each kept when faster; the state is
        restored after every trial."""
        sch, seg, grid, n = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        check(_lib.load().orion_hip_tune(self.ctx, int(steps), C.byref(sch), C.byref(seg), C.byref(grid),
                                         C.byref(ms), C.byref(n)), "tune")
        return {"schedule": self.SCHEDULES[sch.value], "visc_form": "fused" if self.visc_form() else "face_queue",
                "visc_segment": seg.value, "visc_grid": grid.value,
                "ms_per_step": ms.value, "trials": n.value, "steps_per_trial": int(steps)}

    SCHEDULES = ("diagonal", "critical_path", "chained")

    def set_schedule(self, which: int) -> None:
        """The LU-SGS tile launch order (0 diagonal, 1 critical path first,
        the default, both a launch per sweep; 2 the four sweeps chained in
        one launch); the results are the same under every order."""
        check(_lib.load().orion_hip_set_schedule(self.ctx, int(which)), "set_schedule")

    def schedule(self) -> int:
        return int(_lib.load().orion_hip_get_schedule(self.ctx))

    def tune_schedule(self, steps: int = 3):
        """Time ``steps`` pseudo-steps under each schedule from the current
        state (which is restored), keep the fastest; returns (index,
        ms per step of each schedule)."""
        chosen = C.c_int()
        ms = (C.c_double * len(self.SCHEDULES))()
        check(_lib.load().orion_hip_tune_schedule(self.ctx, int(steps), C.byref(chosen), ms), "tune_schedule")
        return int(chosen.value), list(ms)

    # ---------------------------------------------------------------- output
    def output_fields(self, b: int) -> np.ndarray:
        """The VTS point fields of block b, formed on the device as float32:
        Density, Pressure, Temperature, Mach (n interior nodes each, i
        fastest), then Velocity (u, v, w per node); 7 n values
        (PostProcess.cpp:183-240)."""
        i, j, k = (int(v) for v in self.case.dims[b])
        out = np.empty(7 * i * j * k, dtype=np.float32)
        check(_lib.load().orion_hip_output_fields(self.ctx, b, out.ctypes.data_as(C.POINTER(C.c_float)), out.size),
              "output_fields")
        return out

    def save_state(self, path: str, step: int = 0) -> None:
        """Write the restart state (every block's padded prim, after a header
        with the block sizes and ``step``)."""
        check(_lib.load().orion_hip_save_state(self.ctx, str(path).encode(), int(step)), "save_state")

    def load_state(self, path: str) -> int:
        """Restore a state written by save_state for the same grid; returns
        its step."""
        st = C.c_longlong()
        check(_lib.load().orion_hip_load_state(self.ctx, str(path).encode(), C.byref(st)), "load_state")
        return int(st.value)


# ---------------------------------------------------------------------------
# operators (same names as the reference's classes)
# ---------------------------------------------------------------------------
class HaloExchanger:
    def exchange_bc(self, bc, fs: FlowFieldSet) -> None:
        check(_lib.load().orion_hip_exchange_bc(fs.ctx), "exchange_bc")

    def average_interface_residuals(self, bc, fs: FlowFieldSet) -> None:
        check(_lib.load().orion_hip_average_interface_residuals(fs.ctx), "average_interface_residuals")


def apply_physical_bc(bc, fs: FlowFieldSet, params) -> None:
    check(_lib.load().orion_hip_apply_physical_bc(fs.ctx), "apply_physical_bc")


class StateUpdater:
    @staticmethod
    def update_flow_states(fs: FlowFieldSet, params) -> None:
        check(_lib.load().orion_hip_update_flow_states(fs.ctx), "update_flow_states")

    @staticmethod
    def update_solution(fs: FlowFieldSet, bc, params) -> None:
        check(_lib.load().orion_hip_update_solution(fs.ctx), "update_solution")


class TimeIntegrator:
    @staticmethod
    def calculate_time_step(fs: FlowFieldSet, params) -> None:
        v = C.c_double()
        check(_lib.load().orion_hip_calculate_time_step(fs.ctx, C.byref(v)), "calculate_time_step")
        if params is not None:
            params.step.phydtime = v.value


class FluxComputer:
    @staticmethod
    def compute_viscous_rhs(fs: FlowFieldSet, params) -> None:
        check(_lib.load().orion_hip_compute_viscous_rhs(fs.ctx), "compute_viscous_rhs")


class InviscidFluxComputer:
    @staticmethod
    def compute_inviscid_rhs(fs: FlowFieldSet, params) -> None:
        check(_lib.load().orion_hip_compute_inviscid_rhs(fs.ctx), "compute_inviscid_rhs")


class ResidualMonitor:
    @staticmethod
    def compute(fs: FlowFieldSet, params) -> ResidualStatsPy:
        s = ResidualStats()
        check(_lib.load().orion_hip_residual(fs.ctx, C.byref(s)), "residual")
        return _stats(s)


class PostProcess:
    """postprocess::PostProcess (src/postprocess/PostProcess.hpp:10-33)."""

    @staticmethod
    def write_solution(fs: FlowFieldSet, grid, params, step: int) -> None:
        """PostProcess::write_solution (src/postprocess/PostProcess.cpp:44-123):
        one VTS file per block and the VTM index, named from
        params.filename.flowname.  ``grid`` is the case (its grid
        coordinates are the files' points)."""
        flowname = getattr(getattr(params, "filename", None), "flowname", "") or ""
        xyz = np.ascontiguousarray(grid.xyz_flat(), dtype=np.float64)
        check(_lib.load().orion_hip_write_solution(fs.ctx, flowname.encode(), int(step),
                                                   xyz.ctypes.data_as(C.POINTER(C.c_double))), "write_solution")


class OrionSolver:
    """The step loop of OrionSolver::solve (src/core/OrionSolver.cpp:139-209)
    on the device; ``solve`` issues all steps asynchronously and reads the
    residual history back once."""

    def __init__(self, case, keep_q: bool = False, device: int = 0):
        self.case = case
        self.params = make_params(case)
        self.fs = FlowFieldSet(case, keep_q=keep_q, device=device)

    def solve(self, nsteps: Optional[int] = None) -> List[ResidualStatsPy]:
        n = int(nsteps if nsteps is not None else self.case.nsteps)
        hist = (ResidualStats * max(1, n))()
        check(_lib.load().orion_hip_run(self.fs.ctx, n, hist), "run")
        out = [_stats(hist[i]) for i in range(n)]
        if out:
            self.params.step.phydtime = out[-1].phydtime
        return out

    def step_by_operators(self) -> ResidualStatsPy:
        """One step through the individual operators (the reference's call
        sequence), for parity checks of each stage."""
        fs, p, bc = self.fs, self.params, self.case.regions
        halo = HaloExchanger()
        halo.exchange_bc(bc, fs)
        apply_physical_bc(bc, fs, p)
        StateUpdater.update_flow_states(fs, p)
        TimeIntegrator.calculate_time_step(fs, p)
        if p.flowtype.nvis == 1:
            FluxComputer.compute_viscous_rhs(fs, p)
            halo.average_interface_residuals(bc, fs)
        InviscidFluxComputer.compute_inviscid_rhs(fs, p)
        halo.average_interface_residuals(bc, fs)
        StateUpdater.update_solution(fs, bc, p)
        return ResidualMonitor.compute(fs, p)
