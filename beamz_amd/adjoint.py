"""Hand-written adjoint of the 2D (TMxy) time-stepping loop, with
checkpointing, behind a ``torch.autograd.Function``.

The forward run is the ordinary FP32 HIP program: the design density rho
sets the permittivity of an interior box, eps = eps_min + rho (eps_max -
eps_min).  The objective is the mode-overlap power of an Ez line monitor
(a ``FluxMonitor`` with ``components=("Ez",)``):

    O_f = sum_p m_p * measure * DFT_f[Ez](p),     J = sum_f |O_f|^2

with the monitor's own weighted gather and FP64 DFT (``observe.py:529-585``).

Gradient: the exact discrete adjoint of every step (``csrc/adjoint.hip``),
run backward in time.  The forward pass stores the full state every
``checkpoint`` steps; the backward pass walks the segments from the last,
recomputes each segment forward while recording the curl of H over the
design box, then reverses the segment.  dJ/dsE accumulates
sum_n g_n * curl_n (g the adjoint of Ez after step n), and the chain rule
through sE = dt / (eps0 eps_site) and the cell-ownership sampling of Ez
sites (``lattice.py:919-951``: an interior Ez node samples its own cell)
gives dJ/drho.

The reference has no end-to-end FDTD adjoint (its ``execute.py:549-552``
leaves adjoint checkpointing open); its gradient is the host overlap of two
forward runs (``optimization/topology.py:314-342``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _abi as abi
from . import native, runtime
from .api import Simulation
from .boundaries import PML
from .monitors import COMPONENTS


def _clone_state(state):
    return state._replace(**{k: (tuple(t.clone() for t in v) if isinstance(v, tuple) else v.clone())
                             for k, v in state._asdict().items()})


def _copy_state(dst, src):
    for k, v in src._asdict().items():
        d = getattr(dst, k)
        if isinstance(v, tuple):
            for a, b in zip(d, v):
                a.copy_(b)
        else:
            d.copy_(v)


@dataclass
class ForwardRun:
    """What the backward pass needs from a forward run."""

    rho: np.ndarray
    sim: Simulation
    prog: object
    checkpoints: list
    overlaps: np.ndarray      # complex O_f
    J: float


class AdjointProblem2D:
    """A 2D inverse-design objective with its adjoint gradient.

    Parameters
    ----------
    base_eps : (ny, nx) array
        Permittivity outside the design box.
    design : (y0, y1, x0, x1)
        The design box in cells; it must keep clear of the CPML slabs.
    eps_min, eps_max : float
        rho = 0 and rho = 1 permittivities.
    resolution, dt, steps : grid step, time step and run length.
    sources : sources outside the design box (soft, custom or mode sources;
        their coefficients then do not depend on rho).
    monitor : ``FluxMonitor`` with ``components=("Ez",)`` over the output line.
    mode : (P,) array of overlap weights m_p over the monitor's points.
    pml_cells : CPML thickness on every edge.
    checkpoint : steps between stored forward states.
    """

    def __init__(self, base_eps, design, *, eps_min, eps_max, resolution, dt, steps, sources,
                 monitor, mode, pml_cells=10, checkpoint=200):
        self.base_eps = np.asarray(base_eps, dtype=np.float64)
        if self.base_eps.ndim != 2:
            raise ValueError("the adjoint problem is 2D")
        self.design = tuple(int(v) for v in design)
        y0, y1, x0, x1 = self.design
        ny, nx = self.base_eps.shape
        m = max(int(pml_cells), 1)
        if not (m <= y0 < y1 <= ny - m and m <= x0 < x1 <= nx - m):
            raise ValueError("the design box must lie inside the CPML slabs")
        if tuple(monitor.components or ()) != ("Ez",):
            raise ValueError("the objective monitor records Ez only")
        self.eps_min, self.eps_max = float(eps_min), float(eps_max)
        self.resolution, self.dt, self.steps = float(resolution), float(dt), int(steps)
        self.sources, self.monitor = tuple(sources), monitor
        self.mode = np.asarray(mode, dtype=np.float64)
        self.pml_cells, self.checkpoint = int(pml_cells), int(checkpoint)

    # -- forward ---------------------------------------------------------------
    def permittivity(self, rho):
        y0, y1, x0, x1 = self.design
        rho = np.asarray(rho, dtype=np.float64)
        if rho.shape != (y1 - y0, x1 - x0):
            raise ValueError(f"rho must have the design shape {(y1 - y0, x1 - x0)}")
        eps = self.base_eps.copy()
        eps[y0:y1, x0:x1] = self.eps_min + rho * (self.eps_max - self.eps_min)
        return eps

    def simulation(self, rho):
        return Simulation(self.permittivity(rho), self.resolution, dt=self.dt,
                          boundaries=[PML(cells=self.pml_cells)], sources=self.sources,
                          monitors=[self.monitor], steps=self.steps)

    def _overlaps(self, prog, state):
        m = prog.monitors[0]
        F, P = m.freqs.size, m.points
        if self.mode.shape != (P,):
            raise ValueError(f"mode weights must have the monitor's {P} points")
        c = COMPONENTS.index("Ez")
        sl = slice(m.value_offset + c * F * P, m.value_offset + (c + 1) * F * P)
        re = state.dft_re[sl].double().cpu().numpy().reshape(F, P)
        im = state.dft_im[sl].double().cpu().numpy().reshape(F, P)
        return ((re + 1j * im) * (self.mode * m.measure)[None, :]).sum(axis=1)

    def forward(self, rho, keep=True):
        """Run the forward problem; returns (J, ForwardRun)."""
        sim = self.simulation(rho)
        prog = sim.compile()
        state = sim.initial_state()
        checkpoints = []
        done = 0
        while done < self.steps:
            n = min(self.checkpoint, self.steps - done)
            if keep:
                checkpoints.append(_clone_state(state))
            state = runtime.advance(state, prog, n)
            done += n
        torch.cuda.synchronize()
        O = self._overlaps(prog, state)
        J = float(np.sum(np.abs(O) ** 2))
        return J, ForwardRun(np.asarray(rho, dtype=np.float64), sim, prog, checkpoints, O, J)

    def objective(self, rho):
        return self.forward(rho, keep=False)[0]

    # -- adjoint ---------------------------------------------------------------
    def _adjoint_source(self, fwd):
        """Ez-shaped weights W and per-step factors c_n of dJ/dEz_n = c_n W."""
        prog = fwd.prog
        m = prog.monitors[0]
        idx, w = m.plans["Ez"]
        W = np.zeros(int(np.prod(prog.shapes["Ez"])))
        np.add.at(W, idx.reshape(-1), (w * (self.mode * m.measure)[:, None]).reshape(-1))
        n = np.arange(self.steps)
        t = prog.t0 + (n + 1) * prog.dt
        on = (t >= m.t_start) & (t <= m.t_end) & (n % max(m.interval, 1) == 0)
        phase = 2.0 * np.pi * m.freqs[:, None] * t[None, :]
        c = 2.0 * (fwd.overlaps.real[:, None] * np.cos(phase) +
                   fwd.overlaps.imag[:, None] * np.sin(phase)).sum(axis=0)
        c = np.where(on, c, 0.0)
        return W.reshape(prog.shapes["Ez"]), c

    def gradient(self, fwd: ForwardRun):
        """dJ/drho over the design box (float64 numpy array)."""
        prog, dev = fwd.prog, fwd.prog.device
        y0, y1, x0, x1 = self.design
        dh, dw = y1 - y0, x1 - x0
        K = self.checkpoint
        shapes = prog.shapes
        W, c = self._adjoint_source(fwd)
        f32 = dict(dtype=torch.float32, device=dev)
        a_ez = torch.zeros(shapes["Ez"], **f32)
        a_hx = torch.zeros(shapes["Hx"], **f32)
        a_hy = torch.zeros(shapes["Hy"], **f32)
        a_psi = [torch.zeros_like(p) for p in fwd.checkpoints[0].psi_h + fwd.checkpoints[0].psi_e]
        work = [torch.zeros(shapes["Ez"], **f32), torch.zeros(shapes["Ez"], **f32),
                torch.zeros(shapes["Hy"], **f32), torch.zeros(shapes["Hx"], **f32)]
        W_dev = torch.as_tensor(W, **f32)
        c_dev = torch.as_tensor(c, **f32)
        record = torch.zeros((K, dh, dw), **f32)
        grad = torch.zeros((dh, dw), dtype=torch.float64, device=dev)

        adj = native.Adjoint2d()
        co = prog.coeffs
        adj.e_decay = native.buffer(co["e_decay_z"])
        adj.e_source = native.buffer(co["e_source_z"])
        adj.h_decay = native.buffer(co["h_decay_x"])
        adj.h_source = native.buffer(co["h_source_x"])
        adj.cpml_table = native.buffer(prog.cpml_table)
        adj.cpml_cells = int(prog.cpml.cells) if prog.cpml is not None else 0
        adj.metallic = int(prog.boundary_code & 0xFF)
        adj.resolution = float(prog.abi_resolution)
        adj.a_ez, adj.a_hx, adj.a_hy = native.buffer(a_ez), native.buffer(a_hx), native.buffer(a_hy)
        for t in range(4):
            adj.a_psi[t] = native.buffer(a_psi[t])
            adj.work[t] = native.buffer(work[t])
        adj.source_profile = native.buffer(W_dev)
        adj.source_time = native.buffer(c_dev)
        for q, v in enumerate(self.design):
            adj.design[q] = v
        adj.record = native.buffer(record)
        adj.grad = native.buffer(grad)

        state = _clone_state(fwd.checkpoints[0])
        # The recompute steps one state one step at a time: its program call
        # and the record call are described once (the segments restore the
        # checkpoints into the same tensors).  Building the program call anew
        # took ~170 us of host time per step, several times the step's
        # kernels on a 600 x 600 grid.
        step = runtime.prepare_chunk(state, prog, 1)
        rec = native.PreparedAdjointRecord(adj, state.hx, state.hy) if step is not None else None
        nseg = len(fwd.checkpoints)
        for seg in range(nseg - 1, -1, -1):
            first = seg * K
            count = min(K, self.steps - first)
            _copy_state(state, fwd.checkpoints[seg])
            for k in range(count):
                # one forward step of the segment, then record the curl of H'
                if step is not None:
                    step()
                else:
                    state = runtime.run_chunk(state, prog, 1)
                state.current_step.fill_(first + k + 1)
                state.t.fill_(prog.t0 + (first + k + 1) * prog.dt)
                if rec is not None:
                    rec(k)
                else:
                    native.call_adjoint_record(adj, state.hx, state.hy, k)
            native.call_adjoint_reverse(adj, first + count - 1, count, count - 1,
                                        abi.FLAG_GRAPH_CACHE)
        torch.cuda.synchronize()
        g_se = grad.cpu().numpy()
        # sE = dt / (eps0 eps_site) (sigma = 0): dsE/deps = -sE / eps; an
        # interior Ez node samples its own cell (OFFSETS_TM["Ez"] = (0, 0))
        s_e = co["e_source_z"].double().cpu().numpy()[y0:y1, x0:x1] if co["e_source_z"].dim() \
            else float(co["e_source_z"])
        eps = fwd.sim.permittivity[y0:y1, x0:x1]
        return g_se * (-s_e / eps) * (self.eps_max - self.eps_min)


class AdjointObjective(torch.autograd.Function):
    """``J = AdjointObjective.apply(rho, problem)``: forward FDTD, and on
    ``backward`` the checkpointed reverse-time adjoint."""

    @staticmethod
    def forward(ctx, rho, problem):
        J, fwd = problem.forward(rho.detach().double().cpu().numpy())
        ctx.problem, ctx.fwd = problem, fwd
        ctx.rho_dtype, ctx.rho_device = rho.dtype, rho.device
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.problem.gradient(ctx.fwd)
        ctx.fwd = None
        return torch.as_tensor(g * float(grad_out), dtype=ctx.rho_dtype, device=ctx.rho_device), None


def objective(rho, problem):
    return AdjointObjective.apply(rho, problem)
