"""``native.PreparedProgram`` and ``native.PreparedAdjointRecord`` re-issue
exactly the calls ``native.call_program`` and ``native.call_adjoint_record``
make (the same entry point, buffer descriptions, attributes and slot), with
the buffer descriptions built once.  Checked on the CPU against a stub in place
of the library: no compute call is made."""

from __future__ import annotations

import ctypes

import torch

from beamz_amd import native


def _desc(b):
    return (b.data, b.rank, b.element_type, tuple(b.dims))


class _Stub:
    def __init__(self):
        self.calls = []

    def bzh_program(self, s, arr, n, attrs):
        self.calls.append((s.value, [_desc(arr[i]) for i in range(n)], bytes(attrs._obj),
                           ctypes.addressof(arr)))
        return 0

    def bzh_adjoint2d_record(self, s, adj, bx, by, slot):
        self.calls.append((s.value, ctypes.addressof(adj._obj), _desc(bx._obj), _desc(by._obj),
                           slot))
        return 0


def test_prepared_program_reissues_the_same_call(monkeypatch):
    stub = _Stub()
    monkeypatch.setattr(native, "library", lambda: stub)
    args = (torch.zeros(3, 4), torch.ones(5, dtype=torch.float64), None,
            torch.zeros(2, 2, 2, dtype=torch.int32))
    attrs = native.ProgramAttrs()
    attrs.nsteps = 1
    attrs.flags = 7
    s = ctypes.c_void_p(0)
    native.call_program(args, attrs, stream=s)
    prepared = native.PreparedProgram(args, attrs)
    prepared(stream=s)
    prepared(stream=s)
    direct, first, second = stub.calls
    assert direct[1][0] == (args[0].data_ptr(), 2, direct[1][0][2], (3, 4, 0, 0))
    assert direct[1][2][0] is None                 # an empty slot
    for call in (first, second):
        assert call[:3] == direct[:3]
    assert first[3] == second[3]                   # one array, built once


def test_prepared_adjoint_record_reissues_the_same_call(monkeypatch):
    stub = _Stub()
    monkeypatch.setattr(native, "library", lambda: stub)
    adj = native.Adjoint2d()
    hx, hy = torch.zeros(6, 7), torch.zeros(7, 6)
    s = ctypes.c_void_p(0)
    rec = native.PreparedAdjointRecord(adj, hx, hy)
    for slot in (0, 3):
        native.call_adjoint_record(adj, hx, hy, slot, stream=s)
        rec(slot, stream=s)
    a0, p0, a3, p3 = stub.calls
    assert p0 == a0 and p3 == a3
    assert a0[2] == (hx.data_ptr(), 2, a0[2][2], (6, 7, 0, 0)) and a3[4] == 3
