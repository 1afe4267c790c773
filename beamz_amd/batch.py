"""Many independent 2D (TMxy) simulations in one persistent launch per step.

The reference runs one simulation per call and vmaps its FFI sequentially
(``beamz/simulation/cuda/runtime.py:85``); parameter sweeps and optimizer
populations therefore pay one launch chain per simulation.  Here every
simulation of a batch is cut into 62-column items, the items of all
simulations form one list, and the waves of one grid take them round-robin
-- one march launch per time step for the whole batch (after a small launch
that adds the step's pre-E sources), uneven grid sizes included.  A chunk
of up to 256 steps is captured once as a hipGraph and replayed.

Each member is lowered exactly like a single simulation (``program.py``):
same coefficients, CPML profiles, PEC faces and pre-E soft sources, so a
member of a batch reproduces the same simulation run alone.  The tile height
and the grid size are chosen by :func:`autotune` on the actual batch.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi as abi
from . import native
from .api import Simulation
from .program import CompiledProgram

TILE_ROWS = (8, 16, 32, 64)
MAX_SOURCES = 4
_TUNED: dict = {}


@dataclass
class Member:
    """Device state of one simulation of the batch (two field banks)."""

    prog: CompiledProgram
    ez: list
    hx: list
    hy: list
    psi: list            # [bank][4]: Hx_y, Hy_x, Ez_x, Ez_y
    src: list            # [(start (y, x), coeff (sy, sx), wave (T,))]


def _member_sources(prog):
    groups = prog.source_groups
    for gi, g in enumerate(groups):
        if g is not None and gi != 2:
            raise NotImplementedError("batched simulations take pre-E Ez sources only")
    out = []
    g = groups[2]
    if g is None:
        return out
    coeff, wave, starts = g
    for i in range(coeff.shape[0]):
        c = coeff[i, 0].contiguous()          # (sy, sx) of the padded slab
        st = starts[i].cpu().numpy()
        out.append(((int(st[1]), int(st[2])), c, wave[i].contiguous()))
    if len(out) > MAX_SOURCES:
        raise NotImplementedError(f"at most {MAX_SOURCES} source slabs per simulation")
    return out


def _clip_slab(start, coeff, prog):
    """Crop a zero-padded slab to the Ez support (padding never moves a slab)."""
    ny1, nx1 = prog.shapes["Ez"]
    y0, x0 = start
    sy = min(coeff.shape[0], ny1 - y0)
    sx = min(coeff.shape[1], nx1 - x0)
    return coeff[:sy, :sx].contiguous()


class BatchedSimulation2D:
    """A batch of 2D ``Simulation``s stepped together.

    Every member must be 2D (TMxy), have either no CPML or the same CPML
    thickness on all four edges, and carry only pre-E Ez sources (soft or
    custom slabs, at most four).  Monitors are not part of the batched path.
    """

    def __init__(self, simulations, *, tile_rows=None, workgroups=0):
        self.simulations = list(simulations)
        if not self.simulations:
            raise ValueError("empty batch")
        self.progs = []
        for s in self.simulations:
            if not isinstance(s, Simulation):
                raise TypeError("members must be beamz_amd.Simulation objects")
            p = s.compile()
            if p.is_3d:
                raise ValueError("batched simulations are 2D (TMxy)")
            if p.monitors:
                raise NotImplementedError("monitors are not part of the batched path")
            if p.field_dtype != torch.float32:
                raise NotImplementedError("batched simulations store fields in float32")
            if p.cpml is not None and (p.boundary_code >> 8) == 0:
                raise NotImplementedError("batched CPML needs one thickness on every edge")
            self.progs.append(p)
        self.tile_rows = tile_rows
        self.workgroups = int(workgroups)

    @property
    def cells(self):
        return sum(p.cells for p in self.progs)

    def signature(self):
        return tuple((p.grid, p.cpml.cells if p.cpml else 0) for p in self.progs)

    # -- state ---------------------------------------------------------------
    def initial_state(self):
        members = []
        for p in self.progs:
            dev = p.device
            ez = [torch.zeros(p.shapes["Ez"], device=dev) for _ in range(2)]
            hx = [torch.zeros(p.shapes["Hx"], device=dev) for _ in range(2)]
            hy = [torch.zeros(p.shapes["Hy"], device=dev) for _ in range(2)]
            psi = [[], []]
            if p.cpml is not None:
                for t in p.cpml.h_terms + p.cpml.e_terms:
                    shape = list(p.shapes[t.component])
                    shape[t.axis] = t.low + t.high
                    for b in (0, 1):
                        psi[b].append(torch.zeros(shape, device=dev))
            src = [(st, _clip_slab(st, c, p), w) for st, c, w in _member_sources(p)]
            members.append(Member(p, ez, hx, hy, psi, src))
        step = torch.zeros((), dtype=torch.int32, device=self.progs[0].device)
        return {"members": members, "step": step}

    def _descriptors(self, state):
        """The members' device descriptors for the current bank order, built
        once per order and state (they hold only pointers and sizes)."""
        cache = state.setdefault("_desc", {})
        key = tuple(m.ez[0].data_ptr() for m in state["members"])
        if key not in cache:
            cache[key] = self._build_descriptors(state)
        return cache[key]

    def _build_descriptors(self, state):
        arr = (native.Sim2d * len(state["members"]))()
        keep = []
        for i, m in enumerate(state["members"]):
            p, d = m.prog, arr[i]
            for b in (0, 1):
                d.ez[b] = native.buffer(m.ez[b])
                d.hx[b] = native.buffer(m.hx[b])
                d.hy[b] = native.buffer(m.hy[b])
            c = p.coeffs
            d.e_decay = native.buffer(c["e_decay_z"])
            d.e_source = native.buffer(c["e_source_z"])
            d.h_decay = native.buffer(c["h_decay_x"])
            d.h_source = native.buffer(c["h_source_x"])
            if p.cpml is not None:
                d.cpml_table = native.buffer(p.cpml_table)
                d.cpml_cells = int(p.cpml.cells)
                for b in (0, 1):
                    for t in range(4):
                        d.psi[b][t] = native.buffer(m.psi[b][t])
            else:
                d.cpml_table = native.buffer(None)
                d.cpml_cells = 0
            d.nsrc = len(m.src)
            for q, (st, coeff, wave) in enumerate(m.src):
                d.src_coeff[q] = native.buffer(coeff)
                d.src_wave[q] = native.buffer(wave)
                d.src_start[q][0], d.src_start[q][1] = st
                keep += [coeff, wave]
            d.metallic = int(p.boundary_code & 0xFF)
            d.resolution = float(p.abi_resolution)
        return arr, keep

    def _swap(self, state):
        for m in state["members"]:
            m.ez.reverse()
            m.hx.reverse()
            m.hy.reverse()
            m.psi.reverse()

    def run_chunk(self, state, nsteps, *, tile_rows=None, workgroups=None, flags=abi.DEFAULT_FLAGS,
                  descriptors=None):
        """One call of ``nsteps`` steps; ``descriptors`` (from
        :meth:`_descriptors` of this state) skips building them again."""
        tr = tile_rows or self.tile_rows or 32
        wg = self.workgroups if workgroups is None else workgroups
        sims, keep = self._descriptors(state) if descriptors is None else descriptors
        attrs = native.BatchAttrs(abi_version=abi.ABI_VERSION, flags=flags, nsteps=int(nsteps),
                                  tile_rows=int(tr), workgroups=int(wg))
        native.call_batch2d(sims, state["step"], attrs)
        del keep
        if nsteps % 2 == 1:
            self._swap(state)
        return state

    def advance(self, state, nsteps, *, chunk=256):
        """Advance every member by ``nsteps`` (chunks of at most ``chunk``)."""
        if self.tile_rows is None:
            autotune(self)
        done = int(state["step"].item())
        remaining = int(nsteps)
        while remaining > 0:
            n = min(chunk, remaining)
            self.run_chunk(state, n)
            done += n
            remaining -= n
            state["step"].fill_(done)
        return state

    def field(self, state, index, name):
        m = state["members"][index]
        return {"ez": m.ez, "hx": m.hx, "hy": m.hy}[name.lower()][0].double().cpu().numpy()


def autotune(batch: BatchedSimulation2D, *, steps=6, candidates=None, verbose=False):
    """Pick the tile height and persistent grid size of ``batch`` by timing
    each candidate for a few steps on a scratch state; the choice is cached
    per batch signature (grid shapes and CPML thickness).  Returns
    ``(tile_rows, workgroups, timings)``."""
    steps = max(2, steps + steps % 2)  # even: the banks end where they start
    key = batch.signature()
    if key in _TUNED:
        batch.tile_rows, batch.workgroups = _TUNED[key][:2]
        return _TUNED[key]
    props = torch.cuda.get_device_properties(batch.progs[0].device)
    cus = int(props.multi_processor_count)
    if candidates is None:
        # (64 and 96 workgroups per CU measured 1.3-1.5 % faster than 32 on the
        # 64-simulation sweep at 8 rows)
        candidates = [(tr, wg * cus) for tr in TILE_ROWS for wg in (8, 16, 32, 64, 96)]
    scratch = batch.initial_state()
    timings = {}
    stream = torch.cuda.current_stream()
    # (an even step count keeps the banks, so one descriptor set serves all)
    desc = batch._descriptors(scratch)
    for tr, wg in candidates:
        # capture and warm with the same step count as the timed call, so that
        # the timed call replays the cached graph; the descriptors are built
        # outside the timed window, which then holds the GPU work only
        batch.run_chunk(scratch, steps, tile_rows=tr, workgroups=wg, descriptors=desc)
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        batch.run_chunk(scratch, steps, tile_rows=tr, workgroups=wg, descriptors=desc)
        stop.record(stream)
        torch.cuda.synchronize()
        timings[(tr, wg)] = start.elapsed_time(stop) / steps
        if verbose:
            print(f"autotune tile_rows={tr} workgroups={wg}: {timings[(tr, wg)]:.4f} ms/step",
                  flush=True)
    best = min(timings, key=timings.get)
    _TUNED[key] = (best[0], best[1], timings)
    batch.tile_rows, batch.workgroups = best
    del scratch
    return _TUNED[key]
