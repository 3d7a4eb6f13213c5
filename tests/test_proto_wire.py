"""The host library's hand-written protobuf encoding of the snapshot
messages against the protobuf runtime's own serializer, byte for byte (no
GPU: the encoders are host code).  The expected messages are built here from
the reference's construction rules: OutputPanels pushes each ring's nodes as
p(i+1,j), p(i+1,j+1), p(i,j+1), p(i,j) with no time (output.cpp:11-32,
simple_simulator.cpp:315-321); OutputSheet sets t and both sheets with num,
rows, cols and the rings p(i,j), p(i,j+1), p(i+1,j+1), p(i+1,j)
(proto_adaptor.h:115-137, simple_simulator.cpp:323-332)."""
import numpy as np

import proto_schema as P
from uvlm_amd.gpu import encode_state


def state(rng, num=2, rows=2, cols=4, wake_rows=3):
    wn = rng.normal(size=(num, rows + 1, cols + 1, 3))
    wg = rng.normal(size=(num, rows, cols))
    kn = rng.normal(size=(num, wake_rows + 1, cols + 1, 3))
    kg = rng.normal(size=(num, wake_rows, cols))
    return wn, wg, kn, kg


def add_ring(vs, nodes, g):
    v = vs.add(gamma=float(g))
    for p in nodes:
        v.nodes.add(x=float(p[0]), y=float(p[1]), z=float(p[2]))


def expected_panels(wn, wg, kn, kg):
    m = P.CLASSES["Snapshot2"]()
    for pos, gam in ((wn, wg), (kn, kg)):
        num, R, Cc = gam.shape
        for n in range(num):
            for i in range(R):
                for j in range(Cc):
                    add_ring(m.vortices, [pos[n, i + 1, j], pos[n, i + 1, j + 1],
                                          pos[n, i, j + 1], pos[n, i, j]], gam[n, i, j])
    return m


def expected_sheets(wn, wg, kn, kg, t):
    m = P.CLASSES["AllVortexSheets"]()
    m.t = t
    for sheet, (pos, gam) in ((m.wing, (wn, wg)), (m.wake, (kn, kg))):
        num, R, Cc = gam.shape
        sheet.num, sheet.rows, sheet.cols = num, R, Cc
        for n in range(num):
            for i in range(R):
                for j in range(Cc):
                    add_ring(sheet.vortices, [pos[n, i, j], pos[n, i, j + 1],
                                              pos[n, i + 1, j + 1], pos[n, i + 1, j]],
                             gam[n, i, j])
    return m


def test_panels_bytes():
    st = state(np.random.default_rng(1))
    got = encode_state("panels", *st)
    assert got == expected_panels(*st).SerializeToString()
    m = P.parse("Snapshot2", got)
    assert not m.HasField("t") and len(m.vortices) == st[1].size + st[3].size


def test_sheets_bytes():
    st = state(np.random.default_rng(2))
    got = encode_state("sheets", *st, t=0.375)
    assert got == expected_sheets(*st, 0.375).SerializeToString()
    m = P.parse("AllVortexSheets", got)
    assert m.t == 0.375 and m.wake.rows == 3


def test_empty_wake_sheet():
    """Step 0 of a run: the wake has no circulation rows yet."""
    wn, wg, _, _ = state(np.random.default_rng(3))
    kn, kg = np.zeros((2, 1, 5, 3)), np.zeros((2, 0, 4))
    got = encode_state("sheets", wn, wg, kn, kg, t=0.0)
    assert got == expected_sheets(wn, wg, kn, kg, 0.0).SerializeToString()
    m = P.parse("AllVortexSheets", got)
    assert m.HasField("t") and m.wake.rows == 0 and len(m.wake.vortices) == 0
