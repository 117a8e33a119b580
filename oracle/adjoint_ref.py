"""Double-precision reference objective of the adjoint workload.

TEST INFRASTRUCTURE ONLY: the product never imports this module; tests use it
as the finite-difference reference for the adjoint gradients.

J = sum_p |F_p|^2, F_p = sum_{n=1..N} V_p(n) / delta_p * exp(-2 pi i f n dT) * dT,
with V_p(n) the probe edge's voltage after step n, run by the oracle's float64
basic engine (FDTD/engine.cpp:110-222 in double) on an operator whose
coefficients depend on the design in double precision: inside a box around
the design region every edge's C and G come from the quarter-cell average of
eps and kappa (FDTD/operator.cpp:1447-1564, here `RefOperator.effmat`) as
C = eps_eff * A_dual / delta, G = kappa_eff * A_dual / delta
(FDTD/operator.cpp:1189-1254), and vv, vi from them by Calc_ECOperatorPos
(FDTD/operator.cpp:956-984), all without rounding to float32.  Outside the
box the coefficients do not depend on the design; they are the reference's
float32 values (UPML rewrites included), widened to double, and so are the
same in every finite-difference run.
"""
from __future__ import annotations

import numpy as np

from .operator_ref import RefOperator
from .oracle import RefEngine


def coeffs64(ref: RefOperator, lo, hi):
    """(vv, vi, ii, iv) as float64 [3][Nx][Ny][Nz]: the float32 operator
    widened to double, with vv and vi recomputed in double on the edges
    lo <= (x, y, z) <= hi (inclusive mesh indices)."""
    vv = ref.vv.astype(np.float64)
    vi = ref.vi.astype(np.float64)
    ii = ref.ii.astype(np.float64)
    iv = ref.iv.astype(np.float64)
    m = ref.mesh
    N = ref.N
    dT = ref.dT
    sl = tuple(slice(int(lo[a]), int(hi[a]) + 1) for a in range(3))
    for ny in range(3):
        nP, nPP = (ny + 1) % 3, (ny + 2) % 3
        e0, e1 = ref.effmat[ny][0], ref.effmat[ny][1]

        def ax(a, k):
            shape = [1, 1, 1]
            shape[k] = N[k]
            return np.asarray(a, np.float64).reshape(shape)
        delta = ax(m.edge_len(ny, False), ny)
        area = ax(m.edge_len(nP, True), nP) * ax(m.edge_len(nPP, True), nPP)
        Cd = np.broadcast_to(np.asarray(e0, np.float64) * area / delta, tuple(N))[sl]
        Gd = np.broadcast_to(np.asarray(e1, np.float64) * area / delta, tuple(N))[sl]
        a = dT * Gd / 2.0 / Cd
        vv[ny][sl] = (1.0 - a) / (1.0 + a)
        vi[ny][sl] = (dT / Cd) / (1.0 + a)
    return vv, vi, ii, iv


def objective_f64(scene, probes, freq, nsteps, lo, hi):
    """(J, dT) of the scene with the double-precision design coefficients on
    the edges lo..hi; probes: list of (component, (x, y, z))."""
    ref = RefOperator(scene).build()
    c64 = coeffs64(ref, lo, hi)
    eng = RefEngine(ref, "basic_f64", coeffs64=c64)
    V, _ = eng.raw_views()
    d = [np.diff(np.asarray(l, np.float64) * scene["delta_unit"]) for l in scene["lines"]]
    div = np.array([d[c][p[c]] for c, p in probes])
    idx = tuple(np.array([[c] + list(p) for c, p in probes]).T)
    F = np.zeros(len(probes), np.complex128)
    for n in range(1, nsteps + 1):
        eng.iterate(1)
        F += V[idx] / div * (np.exp(-2j * np.pi * freq * n * ref.dT) * ref.dT)
    eng.close()
    return float(np.sum(np.abs(F) ** 2)), ref.dT


def objective_f64_job(args):
    """objective_f64(*args), for a process pool (CPU only)."""
    return objective_f64(*args)
