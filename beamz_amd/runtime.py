"""Lower Yee phases and multi-step programs to the HIP library's positional
entry points.

This module mirrors ``beamz/simulation/cuda/runtime.py`` function for function:
``update_h``/``update_e`` (``runtime.py:184-280``, one phase through
``bzh_streamed``), ``run_steps`` (``:685-750``), ``run_source_group_steps``
(``:496-527``) and ``run_program_steps`` (``:621-682``) through
``bzh_program`` with the same positional layouts (``abi_layout.json``).  State
tensors are updated in place (the reference's ``input_output_aliases``) and
the same state object is returned.
"""

from __future__ import annotations

import ctypes

import torch

from . import _abi as abi
from . import native
from .program import CompiledProgram, SimulationState, field_bank

_EMPTY_GROUP = None


def _empty(device):
    return torch.zeros(0, dtype=torch.float32, device=device)


def _metrics(prog, phase):
    """The (z, y, x) derivative metrics of one phase (``cuda/runtime.py:89-94``):
    none on isotropic uniform grids (metric_kind 0)."""
    table = prog.h_metrics if phase == abi.PHASE_H else prog.e_metrics
    if table:
        return tuple(table)
    e = _empty(prog.device)
    return (e, e, e)


def _materials(prog, phase):
    c = prog.coeffs
    p = "h" if phase == abi.PHASE_H else "e"
    return (c[f"{p}_decay_x"], c[f"{p}_decay_y"], c[f"{p}_decay_z"],
            c[f"{p}_source_x"], c[f"{p}_source_y"], c[f"{p}_source_z"])


def _psi_slots(psi, device):
    slots = list(psi)
    while len(slots) < abi.CPML_TERM_COUNT:
        slots.append(_empty(device))
    return tuple(slots)


def _cpml_payload(prog, state, phase):
    if phase == abi.PHASE_H:
        return (*_materials(prog, phase), prog.h_meta, *prog.h_term_arrays,
                *_psi_slots(state.psi_h, prog.device))
    return (*_materials(prog, phase), prog.e_meta, *prog.e_term_arrays,
            *_psi_slots(state.psi_e, prog.device))


def _graph_io(prog: CompiledProgram, state: SimulationState):
    """Positional inputs of the 24-input (Yee) or 74-input (CPML) graph
    (``runtime.py:284-381``)."""
    if prog.cpml_enabled:
        return (*state.fields(), *_cpml_payload(prog, state, abi.PHASE_H),
                *_cpml_payload(prog, state, abi.PHASE_E), *_metrics(prog, abi.PHASE_H),
                *_metrics(prog, abi.PHASE_E))
    return (*state.fields(), *_materials(prog, abi.PHASE_H), *_materials(prog, abi.PHASE_E),
            *_metrics(prog, abi.PHASE_H), *_metrics(prog, abi.PHASE_E))


def _source_group_arguments(prog):
    out = []
    for g in prog.source_groups:
        if g is None:
            dev = prog.device
            out += [torch.zeros((0, 1, 1, 1), device=dev), torch.zeros((0, 1), device=dev),
                    torch.zeros((0, 3), dtype=torch.int32, device=dev)]
        else:
            out += list(g)
    return tuple(out)


def _program_attrs(prog, nsteps, layout, monitor_count=0):
    return native.ProgramAttrs(
        abi_version=abi.ABI_VERSION, flags=prog.flags, nsteps=int(nsteps), dt=prog.dt,
        resolution=prog.abi_resolution, boundary_code=prog.boundary_code,
        metric_kind=prog.metric_kind, program_layout=layout, cpml_enabled=int(prog.cpml_enabled),
        monitor_count=int(monitor_count), coincident_source_group_mask=prog.coincident_mask)


def _phase(prog, state, phase):
    if phase == abi.PHASE_H:
        targets, sources = (state.hx, state.hy, state.hz), (state.ex, state.ey, state.ez)
    else:
        targets, sources = (state.ex, state.ey, state.ez), (state.hx, state.hy, state.hz)
    args = [*targets, *sources, *_materials(prog, phase)]
    if prog.cpml_enabled:
        meta = prog.h_meta if phase == abi.PHASE_H else prog.e_meta
        arrays = prog.h_term_arrays if phase == abi.PHASE_H else prog.e_term_arrays
        psi = state.psi_h if phase == abi.PHASE_H else state.psi_e
        n = prog.nterms
        args += [meta, *arrays[:3 * n], *psi[:n]]
    args += list(_metrics(prog, phase))
    attrs = native.PhaseAttrs(
        abi_version=abi.ABI_VERSION, flags=prog.flags, phase=phase, nterms=prog.nterms,
        dt=prog.dt, resolution=prog.abi_resolution, boundary_code=prog.boundary_code,
        metric_kind=prog.metric_kind)
    native.call_streamed(args, attrs)
    return state


def update_h(state: SimulationState, prog: CompiledProgram) -> SimulationState:
    """Advance the magnetic fields and their CPML memory (``runtime.py:184-231``)."""
    return _phase(prog, state, abi.PHASE_H)


def update_e(state: SimulationState, prog: CompiledProgram) -> SimulationState:
    """Advance the electric fields and their CPML memory (``runtime.py:234-280``)."""
    return _phase(prog, state, abi.PHASE_E)


def uses_two_banks(prog) -> bool:
    """The fused two-bank step applies to 3D programs with at most
    ``MAX_H_SLABS`` H-timed source slabs when the fused flag is set
    (``abi.FLAG_FUSED_STEP``)."""
    from .program import MAX_H_SLABS, h_source_slabs
    return (prog.is_3d and bool(prog.flags & abi.FLAG_FUSED_STEP)
            and h_source_slabs(prog.source_groups) <= MAX_H_SLABS)


def with_workspace(state):
    """Allocate the second field bank on first use (the CPML bank is allocated
    with the state, in the same arena as the first)."""
    if state.fields_b:
        return state
    fields_b = field_bank([tuple(t.shape) for t in state.fields()], state.hx.dtype,
                          state.hx.device)
    psi_b = state.psi_b
    if len(psi_b) != len(state.psi_h) + len(state.psi_e):
        psi_b = tuple(torch.empty_like(t) for t in (*state.psi_h, *state.psi_e))
    return state._replace(fields_b=fields_b, psi_b=psi_b)


def _swap_banks(state, nsteps):
    """After an odd step count the result lives in the workspace bank
    (``runtime.py:393-440``): exchange the bank roles."""
    if nsteps % 2 == 0:
        return state
    f = state.fields_b
    nh = len(state.psi_h)
    return state._replace(hx=f[0], hy=f[1], hz=f[2], ex=f[3], ey=f[4], ez=f[5],
                          psi_h=tuple(state.psi_b[:nh]), psi_e=tuple(state.psi_b[nh:]),
                          fields_b=state.fields(), psi_b=(*state.psi_h, *state.psi_e))


def _temporal_io(prog, state):
    if prog.cpml_enabled:
        return (*_graph_io(prog, state), *state.fields_b, *state.psi_b)
    return (*_graph_io(prog, state), *state.fields_b)


def run_steps(state, prog, nsteps):
    """Source-free, monitor-free steps (``runtime.py:685-750``).  Returns the
    state holding the result (the bank roles swap after an odd count)."""
    if uses_two_banks(prog):
        return run_source_group_steps(state, prog, nsteps)
    layout = abi.PROGRAM_LAYOUT_CPML_IN_PLACE if prog.cpml_enabled else abi.PROGRAM_LAYOUT_YEE_IN_PLACE
    native.call_program(_graph_io(prog, state), _program_attrs(prog, nsteps, layout))
    return state


def run_source_group_steps(state, prog, nsteps, *, noalloc=False):
    """Steps with the nine slab-source groups (``runtime.py:496-527``);
    ``noalloc`` as in :func:`run_program_steps`."""
    if uses_two_banks(prog):
        state = with_workspace(state)
        args = (*_temporal_io(prog, state), *_source_group_arguments(prog), state.current_step)
        native.call_program(args, _program_attrs(prog, nsteps, abi.PROGRAM_LAYOUT_SOURCE_TEMPORAL_CPML),
                            noalloc=noalloc)
        return _swap_banks(state, nsteps)
    args = (*_graph_io(prog, state), *_source_group_arguments(prog), state.current_step)
    native.call_program(args, _program_attrs(prog, nsteps, abi.PROGRAM_LAYOUT_SOURCE_IN_PLACE),
                        noalloc=noalloc)
    return state


def _monitor_tail(state, prog, noalloc):
    """The source groups, packed monitors, DFT arenas and clocks that follow
    the field arguments of a monitor layout."""
    if state.dft_re.dtype != torch.float64:
        raise ValueError("the HIP program accumulates DFT monitors in float64")
    tail = (*_source_group_arguments(prog), *prog.packed_monitors,
            state.dft_re, state.dft_im, state.dft_w, state.t, state.current_step)
    if prog.monitor_boxes is not None and not noalloc:
        tail = (*tail, prog.monitor_boxes)      # host hint: the cells the monitors read
    return tail


def run_program_steps(state, prog, nsteps, *, noalloc=False):
    """Steps with sources and packed DFT monitors (``runtime.py:621-682``).
    ``noalloc`` calls ``bzh_program_noalloc``, which allocates no device
    memory (the state must already hold its workspace banks)."""
    tail = _monitor_tail(state, prog, noalloc)
    if uses_two_banks(prog):
        state = with_workspace(state)
        native.call_program((*_temporal_io(prog, state), *tail),
                            _program_attrs(prog, nsteps, abi.PROGRAM_LAYOUT_MONITOR_TEMPORAL_CPML,
                                           monitor_count=len(prog.monitors)), noalloc=noalloc)
        return _swap_banks(state, nsteps)
    native.call_program((*_graph_io(prog, state), *tail),
                        _program_attrs(prog, nsteps, abi.PROGRAM_LAYOUT_MONITOR_IN_PLACE,
                                       monitor_count=len(prog.monitors)), noalloc=noalloc)
    return state


def run_chunk(state, prog, nsteps):
    """Pick the layout the program needs (``execute.py:595-681``); returns the
    state that holds the result."""
    if prog.monitors:
        return run_program_steps(state, prog, nsteps)
    if any(g is not None for g in prog.source_groups):
        return run_source_group_steps(state, prog, nsteps)
    return run_steps(state, prog, nsteps)


def prepare_chunk(state, prog, nsteps):
    """The program call :func:`run_chunk` makes for ``nsteps`` steps of
    ``state``, as a :class:`native.PreparedProgram` (arguments marshalled
    once), for loops that step one state many times in short chunks.  Only
    the in-place layouts qualify, whose steps leave every tensor in its role;
    None for a program that steps through two banks."""
    if uses_two_banks(prog):
        return None
    if prog.monitors:
        return native.PreparedProgram(
            (*_graph_io(prog, state), *_monitor_tail(state, prog, False)),
            _program_attrs(prog, nsteps, abi.PROGRAM_LAYOUT_MONITOR_IN_PLACE,
                           monitor_count=len(prog.monitors)))
    if any(g is not None for g in prog.source_groups):
        return native.PreparedProgram(
            (*_graph_io(prog, state), *_source_group_arguments(prog), state.current_step),
            _program_attrs(prog, nsteps, abi.PROGRAM_LAYOUT_SOURCE_IN_PLACE))
    layout = abi.PROGRAM_LAYOUT_CPML_IN_PLACE if prog.cpml_enabled else abi.PROGRAM_LAYOUT_YEE_IN_PLACE
    return native.PreparedProgram(_graph_io(prog, state), _program_attrs(prog, nsteps, layout))


GRAPH_MAX_STEPS = 256  # execute.py:64


def advance(state, prog, nsteps, *, chunk=GRAPH_MAX_STEPS):
    """Advance ``nsteps`` in chunks of at most ``chunk`` steps; the clocks are
    recomputed from the run origin after every chunk (``execute.py:622-657``).
    Returns the state holding the result."""
    done = int(state.current_step.item()) if nsteps > 0 else 0
    remaining = int(nsteps)
    while remaining > 0:
        n = min(chunk, remaining)
        state = run_chunk(state, prog, n)
        done += n
        remaining -= n
        state.current_step.fill_(done)
        state.t.fill_(prog.t0 + done * prog.dt)
    return state


def step_kernel_name(prog, pairs=False):
    """The step kernels as rocprofv3 names them; ``pairs`` when the last
    program call ran two steps per launch."""
    if uses_two_banks(prog):
        # one pair launch runs the general and the core bodies
        return "pair_step" if pairs else "fused_interior+fused_edge"
    return "h_phase_3d+e_phase_3d" if prog.is_3d else "h_phase_tm+e_phase_tm"


def time_step_kernel(state, prog, *, launches=20, stream=None):
    """Average device time of one bare field step (no sources or monitors),
    measured with events on ``stream``: four steps per call (two pair launches
    where the pair kernel runs) so the banks return to their starting roles."""
    stream = stream or torch.cuda.current_stream()
    with torch.cuda.stream(stream):
        layout_state = state
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        if uses_two_banks(prog):
            layout_state = with_workspace(layout_state)
            args = (*_temporal_io(prog, layout_state), *_source_group_arguments_empty(prog),
                    layout_state.current_step)
            attrs = _program_attrs(prog, 4, abi.PROGRAM_LAYOUT_SOURCE_TEMPORAL_CPML)
            attrs.coincident_source_group_mask = 0
        else:
            layout = abi.PROGRAM_LAYOUT_CPML_IN_PLACE if prog.cpml_enabled else abi.PROGRAM_LAYOUT_YEE_IN_PLACE
            args = _graph_io(prog, layout_state)
            attrs = _program_attrs(prog, 4, layout)
        native.call_program(args, attrs)  # warm (captures the graph)
        start.record(stream)
        for _ in range(launches):
            native.call_program(args, attrs)
        stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / (4 * launches)


def _source_group_arguments_empty(prog):
    dev = prog.device
    out = []
    for _ in range(abi.SOURCE_GROUP_COUNT):
        out += [torch.zeros((0, 1, 1, 1), device=dev), torch.zeros((0, 1), device=dev),
                torch.zeros((0, 3), dtype=torch.int32, device=dev)]
    return tuple(out)


# ---- descriptor-level entry points (cuda/src/launch.h:12-77) -----------------
# The same work as the positional calls above, described by the structs that
# the reference's BeamzLaunchStreamed / BeamzLaunchProgram take
# (update.cu:919-971, program.cu:339-354).

def _fill(slots, tensors):
    for i, t in enumerate(tensors):
        slots[i] = native.buffer(t)


def _phase_launch(prog, phase, targets, sources, psi, out_targets, out_psi):
    """A :class:`native.Launch` of one phase: inputs in the streamed order
    (targets, sources, materials, metadata, 18 term arrays, psi), outputs the
    written targets and psi."""
    ln = native.Launch()
    ln.abi_version = abi.ABI_VERSION
    ln.flags = prog.flags
    ln.phase = phase
    ln.nterms = prog.nterms
    ln.metric_kind = prog.metric_kind
    ln.dt = prog.dt
    ln.resolution = prog.abi_resolution
    ln.inv_resolution = 1.0 / prog.abi_resolution
    ln.metallic_edges = prog.boundary_code & 0xFF
    ln.uniform_cpml_thickness = prog.boundary_code >> 8
    ins = [*targets, *sources, *_materials(prog, phase)]
    if prog.cpml_enabled:
        meta = prog.h_meta if phase == abi.PHASE_H else prog.e_meta
        arrays = prog.h_term_arrays if phase == abi.PHASE_H else prog.e_term_arrays
        ins += [meta, *arrays[:18], *_psi_slots(psi, prog.device)]
    _fill(ln.inputs, ins + [None] * (37 - len(ins)))
    _fill(ln.metrics, _metrics(prog, phase))
    outs = [*out_targets]
    if prog.cpml_enabled:
        outs += list(_psi_slots(out_psi, prog.device))
    _fill(ln.outputs, outs + [None] * (9 - len(outs)))
    return ln


def launch_phase(state, prog, phase):
    """One phase in place through ``bzh_launch_phase``."""
    if phase == abi.PHASE_H:
        t, src, psi = (state.hx, state.hy, state.hz), (state.ex, state.ey, state.ez), state.psi_h
    else:
        t, src, psi = (state.ex, state.ey, state.ez), (state.hx, state.hy, state.hz), state.psi_e
    native.call_launch_phase(_phase_launch(prog, phase, t, src, psi, t, psi))
    return state


def launch_program(state, prog, nsteps):
    """``nsteps`` steps with sources and monitors through
    ``bzh_launch_program`` (two field banks where the fused step applies).
    Returns the state holding the result, as :func:`run_chunk` does."""
    two = uses_two_banks(prog)
    if two:
        state = with_workspace(state)
    nh = len(state.psi_h)
    a_h, a_e = (state.hx, state.hy, state.hz), (state.ex, state.ey, state.ez)
    fb = state.fields_b if two else state.fields()
    b_h, b_e = tuple(fb[0:3]), tuple(fb[3:6])
    pb_h = tuple(state.psi_b[:nh]) if two else state.psi_h
    pb_e = tuple(state.psi_b[nh:]) if two else state.psi_e
    pl = native.ProgramLaunch()
    pl.h_ab = _phase_launch(prog, abi.PHASE_H, a_h, a_e, state.psi_h, b_h, pb_h)
    pl.e_ab = _phase_launch(prog, abi.PHASE_E, a_e, b_h, state.psi_e, b_e, pb_e)
    pl.h_ba = _phase_launch(prog, abi.PHASE_H, b_h, b_e, pb_h, a_h, state.psi_h)
    pl.e_ba = _phase_launch(prog, abi.PHASE_E, b_e, a_h, pb_e, a_e, state.psi_e)
    groups = (native.SourceGroup * abi.SOURCE_GROUP_COUNT)()
    flat = _source_group_arguments(prog)
    for gi in range(abi.SOURCE_GROUP_COUNT):
        g = groups[gi]
        g.coefficients = native.buffer(flat[3 * gi])
        g.waveforms = native.buffer(flat[3 * gi + 1])
        g.starts = native.buffer(flat[3 * gi + 2])
        g.current_step = native.buffer(state.current_step)
        g.component = gi % 3
        g.timing = gi // 3
        g.coincident = (prog.coincident_mask >> gi) & 1
    pl.source_groups = groups
    pl.source_group_count = abi.SOURCE_GROUP_COUNT
    keep = [groups]
    if prog.monitors:
        if state.dft_re.dtype != torch.float64:
            raise ValueError("the HIP program accumulates DFT monitors in float64")
        m = native.DftGroup()
        names = ("indices", "weights", "frequencies", "component_masks", "counts", "codes", "windows")
        for name, t in zip(names, prog.packed_monitors):
            setattr(m, name, native.buffer(t))
        m.dft_re = native.buffer(state.dft_re)
        m.dft_im = native.buffer(state.dft_im)
        m.dft_weight = native.buffer(state.dft_w)
        m.time = native.buffer(state.t)
        m.current_step = native.buffer(state.current_step)
        m.monitor_count = len(prog.monitors)
        pl.monitors = ctypes.pointer(m)
        keep.append(m)
    pl.field_bank_count = 2 if two else 1
    pl.nsteps = int(nsteps)
    native.call_launch_program(pl)
    del keep
    return _swap_banks(state, nsteps) if two else state
