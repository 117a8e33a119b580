"""Scene descriptions for the benchmark and parity workloads.

A scene is a plain dict (mesh lines in drawing units, drawing unit in meters,
materials / metals / lumped elements / excitations as axis-aligned boxes,
boundary conditions, excitation signal, probes and dumps).  It carries the
same information an openEMS XML setup of box primitives carries
(openems.cpp:837-975 for the FDTD section; CSXCAD boxes for the geometry).

Workloads (all seeds fixed):

* ``pec_cavity_64``      64^3 lines, 1 mm, vacuum, PEC, soft E_z Gaussian
                         (f0=0, fc=10 GHz) at the centre, 2000 steps.
* ``upml_bricks_256``    256^3 lines, 0.5 mm core graded x1.2 over 10 cells
                         toward every face, 40 dielectric bricks, 8-cell UPML,
                         50 Ohm lumped port with Gaussian fc=20 GHz, 10000 steps,
                         E-field DFT plane at 5/10/15 GHz.
* ``mur_layers_1024``    1024x1024x768 lines, 0.25 mm, 12 dielectric z-layers,
                         Mur on all faces, 15 GHz sinusoidal E_x sheet, 4000 steps.
* ``batched_sweep``      64 independent scenes with U{48..160}^3 lines,
                         10 bricks each, 6-cell UPML, 4000 steps.
"""
from __future__ import annotations

import numpy as np

EPS0 = 8.85418781762e-12


def graded_lines(n_lines, core, n_grade, ratio):
    """Uniform core with geometrically growing cells toward both ends."""
    n_cells = n_lines - 1
    n_core = n_cells - 2 * n_grade
    if n_core < 1:
        raise ValueError("mesh too small for the grading")
    widths = ([core * ratio ** k for k in range(n_grade, 0, -1)] + [core] * n_core
              + [core * ratio ** k for k in range(1, n_grade + 1)])
    return np.concatenate([[0.0], np.cumsum(widths)])


def _probe(ptype, start, stop, weight=1.0, freqs=(), norm_dir=-1, name=""):
    return {"type": ptype, "start": [float(v) for v in start], "stop": [float(v) for v in stop],
            "weight": float(weight), "freqs": list(freqs), "norm_dir": int(norm_dir), "name": name}


def pec_cavity(n=64, nrts=2000, fc=10e9):
    """Uniform n^3-line PEC box with a soft E_z Gaussian source at the centre."""
    l = np.arange(n, dtype=np.float64)
    c = n // 2
    return {
        "name": f"pec_cavity_{n}",
        "lines": [l, l.copy(), l.copy()],
        "delta_unit": 1e-3,
        "bc": [0] * 6,
        "signal": {"type": "gauss", "f0": 0.0, "fc": fc},
        "excitations": [{"type": 0, "val": (0.0, 0.0, 1.0), "start": (c, c, c), "stop": (c, c, c + 1), "priority": 0}],
        "nrts": nrts,
        "oversampling": 4,
        "probes": [
            _probe(2, (20 * n // 64, c, c), (20 * n // 64, c, c), name="et"),
            _probe(0, (40 * n // 64, 40 * n // 64, 0), (40 * n // 64, 40 * n // 64, n - 1), name="ut"),
        ],
    }


def _bricks(rng, lines, count, lo_idx, hi_idx, side, eps_range, kappa_range=None):
    mats = []
    for k in range(count):
        lo, hi, sz = [], [], []
        for a in range(3):
            s = int(rng.integers(side[0], side[1] + 1))
            s = min(s, hi_idx[a] - lo_idx[a] - 1)
            i0 = int(rng.integers(lo_idx[a], hi_idx[a] - s + 1))
            lo.append(lines[a][i0])
            hi.append(lines[a][i0 + s])
        eps = float(rng.uniform(*eps_range))
        kap = float(rng.uniform(*kappa_range)) if kappa_range else 0.0
        mats.append({"eps": eps, "kappa": kap, "priority": k + 1, "start": tuple(lo), "stop": tuple(hi)})
    return mats


def upml_bricks(n=256, nrts=10000, nbricks=40, seed=20260202, pml=8, dump=True):
    """Graded n^3-line mesh, dielectric bricks, UPML, 50 Ohm lumped port."""
    lines = [graded_lines(n, 0.5, 10, 1.2) for _ in range(3)] if n >= 64 else \
        [graded_lines(n, 0.5, max(1, n // 16), 1.2) for _ in range(3)]
    rng = np.random.default_rng(seed)
    margin = pml + 3
    mats = _bricks(rng, lines, nbricks, [margin] * 3, [n - 1 - margin] * 3,
                   (max(2, 8 * n // 256), max(3, 48 * n // 256)), (1.0, 10.0), (0.0, 0.05))
    c = n // 2
    x0, x1 = lines[0][c - 1], lines[0][c]
    y0, y1 = lines[1][c - 1], lines[1][c]
    z0, z1 = lines[2][c - 2], lines[2][c + 2]
    start, stop = (x0, y0, z0), (x1, y1, z1)
    xm, ym, zm = 0.5 * (x0 + x1), 0.5 * (y0 + y1), 0.5 * (z0 + z1)
    scene = {
        "name": f"upml_bricks_{n}",
        "lines": lines,
        "delta_unit": 1e-3,
        "bc": [3] * 6,
        "pml_size": [pml] * 6,
        "materials": mats,
        # LumpedPort (python/openEMS/ports.py:200-243): R box with caps, soft E excitation
        # of -direction*excite, voltage probe weight -1, current probe along the normal
        "lumped": [{"dir": 2, "R": 50.0, "C": 0.0, "L": 0.0, "caps": True, "type": "parallel",
                    "priority": 0, "start": start, "stop": stop}],
        "excitations": [{"type": 0, "val": (0.0, 0.0, -1.0), "start": start, "stop": stop, "priority": 0}],
        "signal": {"type": "gauss", "f0": 0.0, "fc": 20e9},
        "nrts": nrts,
        "oversampling": 4,
        "probes": [
            _probe(0, (xm, ym, z0), (xm, ym, z1), weight=-1.0, name="port_ut1"),
            _probe(1, (x0, y0, zm), (x1, y1, zm), weight=1.0, norm_dir=2, name="port_it1"),
        ],
        "dumps": [],
    }
    if dump:
        scene["dumps"].append({"type": 0, "start": (lines[0][0], lines[1][0], lines[2][c]),
                               "stop": (lines[0][-1], lines[1][-1], lines[2][c]),
                               "freqs": [5e9, 10e9, 15e9], "name": "Ef_zmid"})
    return scene


def dispersive_rlc(n=40, nrts=300, pml=6):
    """Graded n^3-line mesh with UPML, a Drude box reaching into the UPML, a
    two-order Lorentz box (eps and mue poles), a two-order Debye box (the
    second order's relaxation time below 2 dT, so it stays off), a series RLC element, a
    parallel RLC element on the edges' own capacitance and a parallel LC with
    R = 0, driven by a soft E_z Gaussian (fc = 20 GHz)."""
    lines = [graded_lines(n, 0.5, max(1, n // 16), 1.2) for _ in range(3)]
    L = lines
    c = n // 2

    def box(i0, i1):
        return tuple(L[a][i0[a]] for a in range(3)), tuple(L[a][i1[a]] for a in range(3))

    s_drude, e_drude = box((2, c - 6, c - 6), (c - 8, c + 6, c + 6))
    s_lor, e_lor = box((c + 3, c - 5, c - 4), (c + 9, c + 5, c + 4))
    mats = [
        {"eps": 1.5, "kappa": 0.01, "priority": 2, "start": s_drude, "stop": e_drude,
         "lorentz": [{"eps_plasma": 18e9, "eps_relax": 2e-11}]},
        {"eps": 2.0, "mue": 1.2, "priority": 3, "start": s_lor, "stop": e_lor,
         "lorentz": [{"eps_plasma": 12e9, "eps_relax": 5e-11, "eps_pole": 9e9},
                     {"eps_plasma": (6e9, 7e9, 8e9), "mue_plasma": 10e9, "mue_relax": 4e-11, "mue_pole": 8e9}]},
        {"eps": 4.0, "priority": 1, "start": box((c - 4, c + 4, c - 6), (c + 2, c + 9, c + 3))[0],
         "stop": box((c - 4, c + 4, c - 6), (c + 2, c + 9, c + 3))[1]},
        {"eps": 2.5, "priority": 4, "start": box((c - 12, c - 12, c + 5), (c - 6, c - 7, c + 10))[0],
         "stop": box((c - 12, c - 12, c + 5), (c - 6, c - 7, c + 10))[1],
         "debye": [{"eps_delta": (1.0, 0.8, 0.6), "eps_relax": 2e-11}, {"eps_delta": 1.0, "eps_relax": 1e-12}]},
    ]
    s_ser, e_ser = box((c - 2, c - 2, c - 3), (c - 1, c - 1, c + 1))
    s_par, e_par = box((c + 1, c - 3, c - 2), (c + 2, c - 3, c + 2))
    s_lc, e_lc = box((c - 5, c + 1, c - 1), (c - 5, c + 3, c - 1))
    scene = {
        "name": f"dispersive_rlc_{n}",
        "lines": lines,
        "delta_unit": 1e-3,
        "bc": [3] * 6,
        "pml_size": [pml] * 6,
        "materials": mats,
        "lumped": [
            {"dir": 2, "R": 50.0, "C": 1e-11, "L": 1e-10, "caps": True, "type": "series", "priority": 0,
             "start": s_ser, "stop": e_ser},
            {"dir": 2, "R": 100.0, "C": 0.0, "L": 2e-9, "caps": False, "type": "parallel", "priority": 0,
             "start": s_par, "stop": e_par},
            {"dir": 1, "R": 0.0, "C": 2e-13, "L": 5e-10, "caps": False, "type": "parallel", "priority": 0,
             "start": s_lc, "stop": e_lc},
        ],
        "excitations": [{"type": 0, "val": (0.0, 0.0, 1.0), "start": (L[0][c - 3], L[1][c + 2], L[2][c - 2]),
                         "stop": (L[0][c - 3], L[1][c + 2], L[2][c + 2]), "priority": 0}],
        "signal": {"type": "gauss", "f0": 0.0, "fc": 20e9},
        "nrts": nrts,
        "oversampling": 4,
        "probes": [_probe(0, (L[0][c - 2], L[1][c - 2], L[2][c - 3]), (L[0][c - 2], L[1][c - 2], L[2][c + 1]),
                          name="ser_ut")],
        "dumps": [],
    }
    return scene


def wire_antenna(n=36, nrts=400, pml=6):
    """Graded n^3-line mesh with UPML and wire (curve) primitives: a bent
    dipole of two metal polylines (along z, then diagonal in x-z and in
    x-y-z, off the mesh lines) fed by a soft voltage excitation curve across
    its gap (Gaussian, fc = 20 GHz); a hard excitation curve on a second
    diagonal wire; a dielectric box; the feed voltage as a probe."""
    lines = [graded_lines(n, 0.5, max(1, n // 16), 1.2) for _ in range(3)]
    L = lines
    c = n // 2

    def pt(i, j, k):
        # a point given in fractional line indices (off the lines where not whole)
        return tuple(float(np.interp(v, np.arange(len(L[a])), L[a])) for a, v in enumerate((i, j, k)))

    arm_up = [pt(c, c, c + 1), pt(c, c, c + 6), pt(c + 4.3, c, c + 9.6), pt(c + 6.2, c + 2.7, c + 11.1)]
    arm_dn = [pt(c, c, c - 1), pt(c, c, c - 6), pt(c - 3.6, c - 1.2, c - 9.4)]
    return {
        "name": f"wire_antenna_{n}",
        "lines": lines,
        "delta_unit": 1e-3,
        "bc": [3] * 6,
        "pml_size": [pml] * 6,
        "materials": [{"eps": 3.0, "kappa": 0.005, "priority": 1,
                       "start": pt(c - 8, c + 3, c - 5), "stop": pt(c - 2, c + 9, c + 4)}],
        "metals": [{"curve": arm_up}, {"curve": arm_dn}],
        "excitations": [
            {"type": 0, "val": (0.0, 0.0, 1.0), "curve": [pt(c, c, c - 1), pt(c, c, c + 1)]},
            {"type": 1, "val": (0.3, 0.2, 0.1), "delay": 2e-11,
             "curve": [pt(c - 9.2, c - 7.5, c + 3.3), pt(c - 5.1, c - 4.4, c + 6.8)]},
        ],
        "signal": {"type": "gauss", "f0": 0.0, "fc": 20e9},
        "nrts": nrts,
        "oversampling": 4,
        "probes": [_probe(0, pt(c, c, c - 1), pt(c, c, c + 1), name="feed_ut"),
                   _probe(2, pt(c + 5, c - 4, c + 3), pt(c + 5, c - 4, c + 3), name="et")],
        "dumps": [],
    }


def mur_layers(nx=1024, ny=1024, nz=768, nrts=4000, seed=20260203, f0=15e9):
    """Layered dielectric, Mur ABC on every face, sinusoidal E_x sheet source."""
    d = 0.25
    lx = np.arange(nx) * d
    ly = np.arange(ny) * d
    lz = np.arange(nz) * d
    rng = np.random.default_rng(seed)
    mats = []
    zi = 0
    for k in range(12):
        t = int(rng.integers(20, 101))
        eps = float(rng.uniform(1.0, 6.0))
        hi = min(zi + t, nz - 1)
        mats.append({"eps": eps, "priority": k + 1, "start": (-1.0, -1.0, lz[zi]),
                     "stop": (lx[-1] + 1.0, ly[-1] + 1.0, lz[hi])})
        zi = hi
        if zi >= nz - 1:
            break
    zs = min(100, nz // 2)
    probes = []
    for i in range(4):
        for j in range(4):
            px = lx[(i + 1) * nx // 5]
            py = ly[(j + 1) * ny // 5]
            probes.append(_probe(2, (px, py, lz[nz // 2]), (px, py, lz[nz // 2]), name=f"e{i}{j}"))
    probes.append(_probe(0, (lx[nx // 2], ly[ny // 2], lz[0]), (lx[nx // 2], ly[ny // 2], lz[-1]), name="ut"))
    return {
        "name": f"mur_layers_{nx}x{ny}x{nz}",
        "lines": [lx, ly, lz],
        "delta_unit": 1e-3,
        "bc": [2] * 6,
        "materials": mats,
        "excitations": [{"type": 0, "val": (1.0, 0.0, 0.0), "start": (lx[0], ly[0], lz[zs]),
                         "stop": (lx[-1], ly[-1], lz[zs]), "priority": 0}],
        "signal": {"type": "sinus", "f0": f0},
        "nrts": nrts,
        "oversampling": 4,
        "probes": probes,
    }


def batched_sweep(count=64, nrts=4000, seed=20260204, lo=48, hi=160):
    """Independent scenes of uneven size for the batched engine."""
    rng = np.random.default_rng(seed)
    scenes = []
    for k in range(count):
        dims = [int(rng.integers(lo, hi + 1)) for _ in range(3)]
        lines = [np.arange(n) * 0.5 for n in dims]
        sub = np.random.default_rng(int(rng.integers(0, 2 ** 31)))
        mats = _bricks(sub, lines, 10, [9] * 3, [n - 10 for n in dims], (4, max(5, min(dims) // 4)), (1.0, 12.0))
        src = [int(sub.integers(12, n - 13)) for n in dims]
        fc = float(sub.uniform(5e9, 30e9))
        s = (lines[0][src[0]], lines[1][src[1]], lines[2][src[2]])
        e = (lines[0][src[0]], lines[1][src[1]], lines[2][src[2] + 1])
        pz = [int(sub.integers(12, n - 13)) for n in dims]
        scenes.append({
            "name": f"batch_{k}",
            "lines": lines,
            "delta_unit": 1e-3,
            "bc": [3] * 6,
            "pml_size": [6] * 6,
            "materials": mats,
            "excitations": [{"type": 0, "val": (0.0, 0.0, 1.0), "start": s, "stop": e, "priority": 0}],
            "signal": {"type": "gauss", "f0": 0.0, "fc": fc},
            "nrts": nrts,
            "oversampling": 4,
            "probes": [
                _probe(0, (lines[0][src[0]], lines[1][src[1]], lines[2][src[2] - 3]),
                       (lines[0][src[0]], lines[1][src[1]], lines[2][src[2] + 4]), name="ut"),
                _probe(2, (lines[0][pz[0]], lines[1][pz[1]], lines[2][pz[2]]),
                       (lines[0][pz[0]], lines[1][pz[1]], lines[2][pz[2]]), name="et"),
            ],
        })
    return scenes


def adjoint_design(nx=128, ny=128, nz=96, d=0.25, pml=8, design=(48, 48, 8), nrts=8000, seed=20260205,
                   fc=40e9, freq=30e9, eps_range=(1.0, 12.0)):
    """Inverse-design workload: uniform d-mm mesh, UPML, soft E_z Gaussian
    sheet source near the low-x side, a centred design region (material grid,
    eps_r ~ U[eps_range], kappa = 0) and a 4 x 4 set of E_z objective probes
    near the high-x side.  Returns (scene, probes, freq) for
    openems_amd.adjoint.DesignProblem; probes are (component, (x, y, z))."""
    rng = np.random.default_rng(seed)
    lines = [np.arange(n) * d for n in (nx, ny, nz)]
    nd = [int(v) for v in design]
    c0 = [(n - 1 - k) // 2 for n, k in zip((nx, ny, nz), nd)]
    eps = rng.uniform(eps_range[0], eps_range[1], size=(nd[2], nd[1], nd[0]))
    xs = pml + max(2, nx // 16)
    ys0, ys1 = ny // 2 - max(2, ny // 8), ny // 2 + max(2, ny // 8)
    zs0, zs1 = nz // 2 - max(2, nz // 12), nz // 2 + max(2, nz // 12)
    xp = nx - 1 - pml - max(2, nx // 16)
    py = [ny // 2 + k * max(1, ny // 32) for k in (-3, -1, 1, 3)]
    pz = [nz // 2 + k * max(1, nz // 48) for k in (-3, -1, 1, 3)]
    probes = [(2, (xp, y, z)) for z in pz for y in py]
    scene = {
        "name": f"adjoint_design_{nx}x{ny}x{nz}",
        "lines": lines,
        "delta_unit": 1e-3,
        "bc": [3] * 6,
        "pml_size": [pml] * 6,
        "materials": [{"grid": {"cell0": c0, "ncells": nd, "eps": eps, "kappa": np.zeros_like(eps)},
                       "mue": 1.0, "sigma": 0.0, "priority": 10}],
        "excitations": [{"type": 0, "val": (0.0, 0.0, 1.0), "start": (lines[0][xs], lines[1][ys0], lines[2][zs0]),
                         "stop": (lines[0][xs], lines[1][ys1], lines[2][zs1]), "priority": 0}],
        "signal": {"type": "gauss", "f0": 0.0, "fc": fc},
        "nrts": nrts,
        "oversampling": 4,
        "probes": [],
    }
    return scene, probes, freq


WORKLOADS = {
    "pec_cavity_64": lambda: pec_cavity(64, 2000),
    "upml_bricks_256": lambda: upml_bricks(256, 10000),
    "mur_layers_1024": lambda: mur_layers(),
}
