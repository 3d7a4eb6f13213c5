"""Workload definitions (plain data) of the measured configurations.

Each case is a dict understood by ``uvlm_amd.gpu.case_spec``:
``wings`` (generator, digit, chord, span, origin, alpha, plug/flap laws
``(amp, omega, phase)``, Fourier twist/bend), ``rows``/``cols`` rings per
wing, ``forward_flight``, ``dt``, ``advect`` (AdvectFactory name),
``steps``.  Defaults not fixed by the workload statements follow the
reference drivers (multiple_flapping_wings_main.cpp:40-49,
plunging_validation.cpp:21-52): chord 1, U = 1, forward flight (-U, 0, 0),
alpha 5 deg, AR 6, 40 steps per cycle, omega = 2 U k / c.
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed: int):
    """SplitMix64 stream of uint64 (Steele, Lea, Flood 2014)."""
    s = seed & MASK64
    while True:
        s = (s + 0x9E3779B97F4A7C15) & MASK64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def uniform_stream(seed: int):
    g = splitmix64(seed)
    while True:
        yield (next(g) >> 11) * (1.0 / (1 << 53))


ALPHA5 = 5.0 * math.pi / 180.0


def _wing(**kw):
    w = {"generator": "rect", "digit": 0, "chord": 1.0, "span": 4.0,
         "origin": (0.0, 0.0, 0.0), "alpha": ALPHA5,
         "plug": (0.0, 0.0, 0.0), "flap": (0.0, 0.0, 0.0)}
    w.update(kw)
    return w


def plate_c1(steps: int = 200, rows: int = 8, cols: int = 16) -> dict:
    """Flat plate c=1, b=4, 8x16 rings, 5 deg, plunge 0.1c cos(wt), k=0.5,
    dt = c/(8U), Euler."""
    k = 0.5
    omega = 2 * 1.0 * k / 1.0
    return {"name": "plate_8x16_euler", "wings": [_wing(plug=(0.1, omega, 0.0))],
            "rows": rows, "cols": cols, "forward_flight": (-1.0, 0.0, 0.0),
            "dt": 1.0 / 8.0, "advect": "euler", "steps": steps}


def plate_c1_core(steps: int = 200, rows: int = 8, cols: int = 16) -> dict:
    """The plate workload with the configured core cut-off 1e-6 (the
    reference's rule with eps = 1e-6 instead of its hard-coded 1e-10)."""
    c = plate_c1(steps=steps, rows=rows, cols=cols)
    c["name"] = "plate_8x16_euler_core1e-6"
    c["cutoff"] = 1e-6
    return c


def naca4412_c2(steps: int = 1000, rows: int = 20, cols: int = 80) -> dict:
    """NACA4412 camber, AR 6, 20x80 rings, root-hinged flap +/-30 deg at
    k=0.3, 5 deg pitch, 40 steps per cycle, reference "RK4"."""
    k = 0.3
    omega = 2 * 1.0 * k / 1.0
    dt = 2.0 * math.pi / omega / 40
    return {"name": "naca4412_20x80_rk4",
            "wings": [_wing(generator="naca4", digit=44, span=6.0,
                            flap=(math.pi / 6, omega, 0.0))],
            "rows": rows, "cols": cols, "forward_flight": (-1.0, 0.0, 0.0),
            "dt": dt, "advect": "RK4", "steps": steps}


def morphing_c3(steps: int = 400, rows: int = 32, cols: int = 128,
                seed: int = 0xC3) -> dict:
    """Flat wing AR 4, 32x128 rings, 5 deg, random smooth twist/camber
    morphing (4 spanwise x 2 temporal Fourier modes, amplitude 0.05c for the
    camber mode and 0.05 rad for twist), k = 0.5, dt = 1/8, Euler."""
    u = uniform_stream(seed)
    omega = 1.0

    def modes(total):
        a = np.array([[2 * next(u) - 1 for _ in range(2)] for _ in range(4)])
        a *= total / np.abs(a).sum()
        ph = np.array([[2 * math.pi * next(u) for _ in range(2)] for _ in range(4)])
        return a, ph

    ta, tp = modes(0.05)
    ba, bp = modes(0.05)
    w = _wing(span=4.0, fourier_omega=omega, twist_amp=ta.tolist(), twist_phase=tp.tolist(),
              bend_amp=ba.tolist(), bend_phase=bp.tolist())
    return {"name": "morphing_32x128_euler", "wings": [w], "rows": rows, "cols": cols,
            "forward_flight": (-1.0, 0.0, 0.0), "dt": 1.0 / 8.0, "advect": "euler",
            "steps": steps}


def batch_c4(cases: int = 256, steps: int = 300, seed: int = 0xC4) -> list:
    """256 independent rectangular wings: R in [4,24], C = 2*U{4..32},
    AR in [2,8], plunge amplitude in [0.05,0.3]c, k in [0.1,1.0],
    40 steps per cycle, RK2."""
    u = uniform_stream(seed)
    out = []
    for i in range(cases):
        R = 4 + int(next(u) * 21)
        C = 2 * (4 + int(next(u) * 29))
        ar = 2 + 6 * next(u)
        amp = 0.05 + 0.25 * next(u)
        k = 0.1 + 0.9 * next(u)
        omega = 2 * k
        out.append({"name": f"batch_{i}", "wings": [_wing(span=ar, plug=(amp, omega, 0.0))],
                    "rows": R, "cols": C, "forward_flight": (-1.0, 0.0, 0.0),
                    "dt": 2 * math.pi / omega / 40, "advect": "RK2", "steps": steps})
    return out


def tandem_c5(steps: int = 2000, rows: int = 16, cols: int = 64) -> dict:
    """Two NACA 8300-camber wings, AR 6, 16x64 rings each, LE-to-LE 2.5c
    (gap 1.5c), flap -(pi/4) cos(wt + n*90deg), k = 0.1, Euler."""
    k = 0.1
    omega = 2 * k
    wings = [_wing(generator="naca4", digit=83, span=6.0, origin=(2.5 * n, 0.0, 0.0),
                   flap=(-math.pi / 4, omega, n * math.pi / 2)) for n in range(2)]
    return {"name": "tandem_2x16x64_euler", "wings": wings, "rows": rows, "cols": cols,
            "forward_flight": (-1.0, 0.0, 0.0), "dt": 2 * math.pi / omega / 40,
            "advect": "euler", "steps": steps}


def pair_model(case: dict, steps: int | None = None, stages: int | None = None) -> float:
    """Biot-Savart pairs of the reference loop over steps 0..steps:
    4 [N^2 + N Nw + L (N + Nw) + S V (N + Nw)] per step."""
    steps = case["steps"] if steps is None else steps
    num = len(case["wings"])
    R, C = case["rows"], case["cols"]
    N = num * R * C
    L = num * (4 * R * C - C)
    S = stages if stages is not None else {"euler": 1, "RK2": 2, "RK4": 4, "AB2": 1,
                                           "RK4classic": 4}[case.get("advect", "euler")]
    tot = 0.0
    for s in range(steps + 1):
        Nw = num * s * C
        V = num * (s + 1) * (C + 1)
        tot += 4.0 * (N * N + N * Nw + L * (N + Nw) + S * V * (N + Nw))
    return tot


WORKLOADS = {"c1": plate_c1, "c2": naca4412_c2, "c3": morphing_c3, "c5": tandem_c5}
