"""Modal S-parameters from DFT monitors: 2D (TMxy) lines and 3D planes.

Restates the reference's extraction for the TM branch of a 2D simulation:
``beamz/analysis/sparameters.py:164-423`` (port waves), ``424-566``
(``s_parameters``: incident/scattered selection, incident floor, power
diagnostics), ``92-108`` (safe ratio, wave selectors),
``beamz/analysis/mode_projection.py:433-546`` (port projection: unit-power
mode, forward/backward mode matrix, pseudo-inverse) and
``beamz/analysis/modal_projection/geometry.py:120-152`` (port components and
flux signs).

Each port names a ``FluxMonitor`` line normal to its axis.  At every
frequency the port's mode is solved on the monitor line (the solver of
``modes.py``), sampled at the monitor points, scaled to unit modal power, and
the monitor's E/H phasors are projected on the forward and backward mode:
``[E; H] = a_plus [e; h] + a_minus [e; -h]``.  This build's line monitors
interpolate E and H onto the same points, so no E-to-H plane delay enters.
"plus" is the wave travelling towards +axis on both axes (the mode's H takes
the physical sign from Faraday's law; the reference's solver flips H for
y-normal planes, ``plane.py:196-197``, which this extraction does not copy).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .constants import LIGHT_SPEED, MU_0
from .modes import solve_line_modes
from .modes3d import _AXIS_POS, _cyclic, _plane_index, _to_bc, solve_plane_modes

# (E component, H component, signed flux sign) of a TM port by axis
_TM_PARTS = {"x": ("Ez", "Hy", -1.0), "y": ("Ez", "Hx", 1.0)}
# (E_b, E_c, H_b, H_c) of a 3D port, (a, b, c) cyclic
_PARTS_3D = {"x": ("Ey", "Ez", "Hy", "Hz"), "y": ("Ez", "Ex", "Hz", "Hx"),
             "z": ("Ex", "Ey", "Hx", "Hy")}


@dataclass(frozen=True)
class Port:
    """A modal port on a monitor line: ``direction`` "+" when the incident
    wave travels along +axis into the device."""

    name: str
    monitor: str
    axis: str
    direction: str = "+"
    mode_index: int = 0
    polarization: str = "tm"

    def __post_init__(self):
        if self.axis not in _PARTS_3D:
            raise ValueError("a port is normal to x, y or z")
        if self.direction not in ("+", "-"):
            raise ValueError("port direction must be '+' or '-'")
        if self.polarization not in ("tm", "vector"):
            raise ValueError("ports are 'tm' (2D TMxy) or 'vector' (3D)")


@dataclass(frozen=True)
class SParameterResult:
    s_matrix: dict
    frequencies: np.ndarray
    diagnostics: dict = field(default_factory=dict)


def wave_selectors(port: Port):
    """(incident, scattered) branch of a 2D port (``sparameters.py:99-107``)."""
    incident = "plus" if port.direction == "+" else "minus"
    return incident, "minus" if incident == "plus" else "plus"


def safe_ratio(num, den, eps=1e-18):
    out = np.zeros_like(np.asarray(num), dtype=np.complex128)
    valid = np.abs(den) > eps
    out[valid] = np.asarray(num)[valid] / np.asarray(den)[valid]
    return out


def port_projection(e_line, frequency, dx, points, axis, mode_index=0):
    """Unit-power forward mode (e, h) at the monitor points and the
    pseudo-inverse of [[e, e], [h, -h]] (``mode_projection.py:433-546``).

    ``e_line`` is the Ez-site permittivity along the monitor line (one value
    per Ez node); ``points`` are the monitor's cell indices along the line
    (sample j sits between nodes j and j + 1)."""
    omega = 2.0 * np.pi * float(frequency)
    eps = np.asarray(e_line, dtype=np.float64)
    target = 0.98 * np.sqrt(max(float(eps.max()), 1e-12))
    m = int(mode_index) + 1
    neffs, ezs = solve_line_modes(eps, omega, dx, 2 * m + 5, target)
    if neffs.size <= mode_index:
        raise ValueError(f"the mode solver found {neffs.size} modes for mode_index={mode_index}")
    neff = float(neffs[mode_index])
    node = ezs[mode_index]
    pts = np.asarray(points, dtype=np.int64)
    e = 0.5 * (node[pts] + node[pts + 1]).astype(np.complex128)
    # transverse H of the mode travelling towards +axis: Faraday's law gives
    # Hy = -beta Ez / (omega mu0) along x and Hx = +beta Ez / (omega mu0)
    # along y, so "plus" is always the wave travelling towards +axis
    sign = _TM_PARTS[axis][2]
    h = sign * neff * omega / LIGHT_SPEED / (omega * MU_0) * e
    power = 0.5 * np.real(np.sum(dx * sign * e * np.conjugate(h)))
    norm = np.sqrt(max(abs(power), 1e-30))
    e, h = e / norm, h / norm
    mat = np.column_stack([np.concatenate([e, h]), np.concatenate([e, -h])])
    return {"e": e, "h": h, "mode_matrix": mat, "pinv": np.linalg.pinv(mat),
            "condition_number": float(np.linalg.cond(mat)), "mode_neff": neff}


def port_waves(results, port: Port, frequencies):
    """(a_plus, a_minus) of ``port`` at ``frequencies`` from the monitor's
    phasors (``sparameters.py:164-423``)."""
    prog = results.program
    idx = next((i for i, m in enumerate(prog.monitors) if m.name == port.monitor), None)
    if idx is None:
        raise ValueError(f"Missing monitor '{port.monitor}' for port '{port.name}'.")
    mon = prog.monitors[idx]
    if mon.normal != port.axis:
        raise ValueError(f"port '{port.name}' axis {port.axis} does not match its monitor")
    is_3d = np.ndim(prog.eps_site["Ez"]) == 3
    if getattr(mon, "placed", False):
        node = placed_port_plane(mon, prog.resolution) if is_3d else None
        if node is None:
            raise ValueError(
                f"port '{port.name}': its monitor '{port.monitor}' is placed by centre and size off "
                "the lattice; a port needs a 3D plane on an E-node plane with its edges on cell "
                "faces, or an index-space monitor (normal, index, ranges)")
        return _port_waves_3d(results, port, frequencies, idx, mon, node)
    if is_3d:
        return _port_waves_3d(results, port, frequencies, idx, mon)
    if port.axis not in _TM_PARTS:
        raise ValueError("a 2D port is normal to x or y")
    e_comp, h_comp, _ = _TM_PARTS[port.axis]
    ph = results.monitor_phasors(idx)
    if ph.get(e_comp) is None or ph.get(h_comp) is None:
        raise ValueError(f"monitor '{port.monitor}' must record {e_comp} and {h_comp}")
    # the monitor's line: Ez-site permittivity along its tangential axis
    eps_ez = np.asarray(prog.eps_site["Ez"])
    line_index = int(mon_index(prog, idx))
    e_line = eps_ez[:, line_index] if port.axis == "x" else eps_ez[line_index, :]
    points = monitor_points(prog, idx)
    out = {"a_plus": np.zeros(len(frequencies), complex), "a_minus": np.zeros(len(frequencies), complex),
           "condition_number": np.zeros(len(frequencies)), "mode_neff": np.zeros(len(frequencies))}
    mfreqs = np.asarray(mon.freqs)
    for k, f in enumerate(frequencies):
        j = int(np.argmin(np.abs(mfreqs - f)))
        if not np.isclose(mfreqs[j], f, rtol=1e-12, atol=0.0):
            raise ValueError(f"monitor '{port.monitor}' holds no DFT at {f} Hz")
        proj = port_projection(e_line, f, prog.resolution, points, port.axis, port.mode_index)
        vec = np.concatenate([ph[e_comp][j], ph[h_comp][j]])
        a = proj["pinv"] @ vec
        out["a_plus"][k], out["a_minus"][k] = a[0], a[1]
        out["condition_number"][k] = proj["condition_number"]
        out["mode_neff"][k] = proj["mode_neff"]
    out["P_plus"] = np.abs(out["a_plus"]) ** 2
    out["P_minus"] = np.abs(out["a_minus"]) ** 2
    return out


def _centres(axis, eb, ec, hb, hc):
    """Mode fields of a Nb x Nc window averaged onto its cell centres, in the
    3D plane monitor's point order (tangential axes in storage order)."""
    eb_q = 0.5 * (eb[:, :-1] + eb[:, 1:])
    hc_q = 0.5 * (hc[:, :-1] + hc[:, 1:])
    ec_q = 0.5 * (ec[:-1, :] + ec[1:, :])
    hb_q = 0.5 * (hb[:-1, :] + hb[1:, :])
    _, b, c = _cyclic(axis)
    flip = _AXIS_POS[b] > _AXIS_POS[c]   # (b, c) not in storage order
    return [v.T.ravel() if flip else v.ravel() for v in (eb_q, ec_q, hb_q, hc_q)]


def port_projection_3d(eps_site, grid, dx, frequency, axis, index, ranges, mode_index=0):
    """Unit-power forward plane mode at a 3D monitor's points and the
    pseudo-inverse of [[e, e], [h, -h]] (``mode_projection.py:162-432``).

    The mode is solved on the monitor's window (``ranges``: tangential cell
    intervals), bounded by PEC, and averaged onto the cell centres where the
    monitor samples E and H; it is renormalized to unit power there."""
    a, b, c = _cyclic(axis)
    rng = dict(ranges)
    (b0, b1), (c0, c1) = rng[b], rng[c]

    def sl(comp_axis_b, comp_axis_c):
        return _plane_index(axis, index, slice(b0, b1 + comp_axis_b), slice(c0, c1 + comp_axis_c))

    eps_b = _to_bc(axis, np.asarray(eps_site["E" + b])[sl(0, 1)])
    eps_c = _to_bc(axis, np.asarray(eps_site["E" + c])[sl(1, 0)])
    ea = np.asarray(eps_site["E" + a])
    lo, hi = max(index - 1, 0), min(index, ea.shape[_AXIS_POS[a]] - 1)
    eps_a = 0.5 * (_to_bc(axis, ea[_plane_index(axis, lo, slice(b0, b1 + 1), slice(c0, c1 + 1))]) +
                   _to_bc(axis, ea[_plane_index(axis, hi, slice(b0, b1 + 1), slice(c0, c1 + 1))]))
    omega = 2.0 * np.pi * float(frequency)
    target = 0.98 * np.sqrt(max(float(eps_b.max()), float(eps_c.max())))
    m = int(mode_index) + 1
    modes = solve_plane_modes(eps_b, eps_c, eps_a, omega, dx, 2 * m + 5, target)
    if modes.neff.size <= mode_index:
        raise ValueError(f"the plane mode solver found {modes.neff.size} modes for mode_index={mode_index}")
    k = int(mode_index)
    eb, ec, hb, hc = _centres(axis, modes.eb[k], modes.ec[k], modes.hb[k], modes.hc[k])
    power = 0.5 * float(np.sum(eb * hc) - np.sum(ec * hb)) * dx * dx
    s = 1.0 / np.sqrt(max(abs(power), 1e-300))
    e = np.concatenate([eb, ec]).astype(np.complex128) * s
    h = np.concatenate([hb, hc]).astype(np.complex128) * s
    mat = np.column_stack([np.concatenate([e, h]), np.concatenate([e, -h])])
    return {"e": e, "h": h, "mode_matrix": mat, "pinv": np.linalg.pinv(mat),
            "condition_number": float(np.linalg.cond(mat)), "mode_neff": float(modes.neff[k])}


def placed_port_plane(mon, dx, tol=1e-6):
    """The E-node plane index of a placed 3D monitor whose samples are those
    of an index-space monitor (normal coordinate on an E-node plane, samples
    on the cell centres of its snapped ranges), or None."""
    node = float(mon.position) / float(dx)
    if abs(node - round(node)) > tol or len(mon.samples) != 2:
        return None
    for (axis, (r0, r1)), coords in zip(mon.ranges, mon.samples):
        want = (np.arange(r0, r1) + 0.5) * dx
        if coords.shape != want.shape or np.max(np.abs(coords - want), initial=0.0) > tol * dx:
            return None
    return int(round(node))


def _port_waves_3d(results, port, frequencies, idx, mon, index=None):
    prog = results.program
    parts = _PARTS_3D[port.axis]
    ph = results.monitor_phasors(idx)
    if any(ph.get(c) is None for c in parts):
        raise ValueError(f"monitor '{port.monitor}' must record {', '.join(parts)}")
    out = {"a_plus": np.zeros(len(frequencies), complex), "a_minus": np.zeros(len(frequencies), complex),
           "condition_number": np.zeros(len(frequencies)), "mode_neff": np.zeros(len(frequencies))}
    mfreqs = np.asarray(mon.freqs)
    for k, f in enumerate(frequencies):
        j = int(np.argmin(np.abs(mfreqs - f)))
        if not np.isclose(mfreqs[j], f, rtol=1e-12, atol=0.0):
            raise ValueError(f"monitor '{port.monitor}' holds no DFT at {f} Hz")
        proj = port_projection_3d(prog.eps_site, prog.grid, prog.resolution, f, port.axis,
                                  mon.index if index is None else index, mon.ranges, port.mode_index)
        vec = np.concatenate([ph[c][j] for c in parts])
        a = proj["pinv"] @ vec
        out["a_plus"][k], out["a_minus"][k] = a[0], a[1]
        out["condition_number"][k] = proj["condition_number"]
        out["mode_neff"][k] = proj["mode_neff"]
    out["P_plus"] = np.abs(out["a_plus"]) ** 2
    out["P_minus"] = np.abs(out["a_minus"]) ** 2
    return out


def mon_index(prog, i):
    """The normal index of monitor ``i`` (kept on the compiled monitor)."""
    return prog.monitors[i].index


def monitor_points(prog, i):
    """Cell indices of monitor ``i``'s samples along its line."""
    return np.arange(*prog.monitors[i].cell_range)


def s_parameters(results, source_port, ports, output_ports=None, frequencies=None,
                 min_incident_db=-40.0):
    """Modal S-parameters ``S[(out, src)] = b_out / a_src`` with power
    diagnostics (``sparameters.py:424-566``)."""
    port_map = {p.name: p for p in ports}
    if len(port_map) != len(ports):
        raise ValueError("port names must be unique")
    src_name = source_port.name if isinstance(source_port, Port) else str(source_port)
    if src_name not in port_map:
        raise ValueError(f"source_port '{src_name}' not found in ports.")
    if frequencies is None:
        mon = next(m for m in results.program.monitors if m.name == port_map[src_name].monitor)
        frequencies = np.asarray(mon.freqs, dtype=float)
    freqs = np.atleast_1d(np.asarray(frequencies, dtype=float))
    if freqs.size == 0 or np.any(freqs <= 0):
        raise ValueError("frequencies must be strictly positive.")
    outs = list(port_map) if output_ports is None else [
        p.name if isinstance(p, Port) else str(p) for p in output_ports]
    waves = {name: port_waves(results, p, freqs) for name, p in port_map.items()}
    inc, _ = wave_selectors(port_map[src_name])
    a_in = waves[src_name][f"a_{inc}"]
    peak = float(np.max(np.abs(a_in))) if a_in.size else 0.0
    floor = max(1e-18, peak * 10.0 ** (float(min_incident_db) / 20.0))
    valid = np.abs(a_in) >= floor
    s = {}
    p_out = np.zeros(freqs.size)
    for name in outs:
        _, scat = wave_selectors(port_map[name])
        b = waves[name][f"a_{scat}"]
        s[(name, src_name)] = np.where(valid, safe_ratio(b, a_in), 0.0 + 0.0j)
        p_out += np.abs(b) ** 2
    p_in = np.abs(a_in) ** 2
    power_sum = p_out / np.maximum(p_in, 1e-18)
    diagnostics = {
        "frequencies": freqs, "source_port": src_name, "output_ports": outs, "waves": waves,
        "P_in": p_in, "P_guided_out": p_out,
        "power_sum": np.where(valid, power_sum, np.nan),
        "loss_est": np.where(valid, 1.0 - power_sum, np.nan), "valid_mask": valid,
    }
    return SParameterResult(s_matrix=s, frequencies=freqs, diagnostics=diagnostics)
