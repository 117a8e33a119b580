"""Benchmark: EC-FDTD cell-updates/s on one MI355X for the 256^3 UPML workload.

A step is one full timestep of the hot path (voltage + current half steps with
the fused UPML, lumped-port source, port probes and the 256x256 E-field DFT
plane at 3 frequencies) on synthetic input: the seeded 40-brick scene of
openems_amd.scenes.upml_bricks.  Fields, coefficients and extension tables are
resident in HBM before timing starts.

Prints one JSON line.  value = prod(numLines) * steps / wall time (MCells/s),
the reference's own speed definition (openems.cpp:1410,1484-1485).

roofline: the dominant (voltage or current update) kernel, achieved =
algorithmic bytes per launch / average launch duration measured with HIP
events on the engine stream over the timed steps.  Algorithmic bytes per cell
update of one half step, by coefficient layout:
  plain      : 15 floats = 60 B (read field, 3 curl inputs read once,
               2x3 coefficients, write field);
  compressed : 9 floats + one uint16 table index = 38 B (coefficients come
               from a table of unique tuples that stays in cache);
plus, per UPML position, 15 floats = 60 B (plain UPML: 3 flux read + 3 flux
write + 9 coefficients) or 6 floats + one uint32 index = 28 B (compressed
UPML).  The single-sweep steps (k_step, k_step2) move a whole timestep per
launch: per cell 12 floats (read and write V and I) + 12 coefficient floats
(plain) or one uint16 index (compressed) = 96 B or 50 B; per UPML position
the two half steps' UPML bytes (120 B plain, 52 B compressed).  `plain_ceiling_frac` compares the cell-update rate with the HBM
ceiling of the plain layout (8 TB/s / 143.57 B per cell-step).

cpu_baseline: the oracle's restated multithreaded SSE-compressed engine
(the reference's default "fastest" engine) on the same operator, on a bounded
number of timesteps on the host's CPU share (its affinity set, capped by
OMP_NUM_THREADS), plus the single-threaded SSE-compressed engine, with the CPU
model.

Beside the headline the line carries the other single-GPU workloads: the 64^3
PEC cavity (plain_cavity), the 1024x1024x768 Mur layers (memory_scale), the
64-instance batched sweep (batched) and the adjoint gradient (adjoint), and
the achievable streaming bandwidth of the device (roofline.achievable).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MCells/s & % HBM roofline (256³ UPML); batched speedup; adjoint/forward time"


LAYOUT_BYTES = {"plain": 60.0, "compressed": 38.0}        # per cell, per half step
PML_BYTES = {"plain": 60.0, "compressed": 28.0}           # per UPML position, per half step
FUSED_BYTES = {"plain": 96.0, "compressed": 50.0}         # per cell, per fused step
FUSED_PML_BYTES = {"plain": 120.0, "compressed": 52.0}    # per UPML position, per fused step
PLAIN_CELL_STEP_BYTES = 120.0


def pml_positions(op):
    return sum(s["num"][0] * s["num"][1] * s["num"][2] for s in op.upml_slabs())


def traffic_from_profile(kernel, layout, size=None, pattern=None):
    """HBM bytes per launch of `kernel` measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs of this workload and coefficient layout (committed summary), or None."""
    import glob
    pattern = pattern or f"upml_bricks_{size}_traffic*.json"
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", pattern))):
        try:
            with open(path) as f:
                d = json.load(f)
        except (OSError, ValueError):
            continue
        if d.get("layout") != layout:
            continue
        for name, k in d.get("kernels", {}).items():
            if name.split("<")[0] == kernel:
                return k["hbm_bytes_per_launch"]
    return None


def cpu_model():
    """Host CPU model name (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def cpu_threads():
    """Threads for the multithreaded CPU engine: the CPUs this process may run
    on (its affinity set), capped by OMP_NUM_THREADS when that is set (the
    host's CPU share for this job)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def _timed_steps(eng, budget_s):
    eng.iterate(2)
    t0 = time.perf_counter()
    eng.iterate(4)
    per = (time.perf_counter() - t0) / 4
    steps = max(4, int(budget_s / max(per, 1e-6)))
    t0 = time.perf_counter()
    eng.iterate(steps)
    return steps, time.perf_counter() - t0


def cpu_baseline(op, nthreads, budget_s=15.0, single_budget_s=8.0):
    """Time the oracle's restated reference engines on this operator: the
    multithreaded SSE-compressed engine (the reference's default) on
    `nthreads` threads, and the single-threaded SSE-compressed engine."""
    import numpy as np
    from oracle.operator_ref import ExcList, UPMLSlab
    from oracle.oracle import RefEngine

    class _Op:
        pass
    ro = _Op()
    ro.N = op.N
    ro.vv, ro.vi, ro.ii, ro.iv = op.coeff("vv"), op.coeff("vi"), op.coeff("ii"), op.coeff("iv")
    ve, ce = op.excitation(False), op.excitation(True)
    ro.vexc = ExcList(ve["index"], ve["dir"], ve["amp"], ve["delay"])
    ro.cexc = ExcList(ce["index"], ce["dir"], ce["amp"], ce["delay"])
    ro.sig_volt, ro.sig_curr = op.signal()
    ro.exc_length = op.signal_length
    ro.period, ro.dT = op.period, op.dT
    ro.mur = []
    ro.upml = []
    for s in op.upml_slabs():
        u = UPMLSlab(s["start"], s["num"])
        for k in ("vv", "vvfo", "vvfn", "ii", "iifo", "iifn"):
            setattr(u, k, s[k])
        ro.upml.append(u)
    cells = float(np.prod(op.N))
    eng = RefEngine(ro, "multithreaded", nthreads)
    steps, dt = _timed_steps(eng, budget_s)
    eng.close()
    out = {"value": round(cells * steps / dt / 1e6, 2), "unit": "MCells/s", "cores": int(nthreads),
           "kind": "port",
           "sample": f"restated multithreaded sse-compressed engine (x-slab threads + barriers, "
                     f"FTZ/DAZ, UPML hooks), {steps} timesteps of the same {op.N[0]}x{op.N[1]}x{op.N[2]} UPML "
                     f"operator, {dt:.1f} s",
           "cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads": int(nthreads)}
    if single_budget_s > 0:
        eng = RefEngine(ro, "sse_compressed", 1)
        s1, d1 = _timed_steps(eng, single_budget_s)
        eng.close()
        out["sse_compressed_1thread"] = {"value": round(cells * s1 / d1 / 1e6, 2), "unit": "MCells/s",
                                         "cores": 1, "sample": f"{s1} timesteps, {d1:.1f} s"}
    return out


def copy_calibration(reps=20, nbytes=1 << 30):
    """A plain device copy of known size (torch's copy kernel, nbytes read + nbytes
    written per launch): the reference point that scales rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                                           "launches": reps}}), flush=True)


def hbm_probe(reads, writes, nontemporal, bytes_per_array, reps=20):
    """GB/s of a streaming kernel reading `reads` and writing `writes` float
    arrays of `bytes_per_array` each, 16 B per lane (libems_hbm_probe.so)."""
    import ctypes as C
    from openems_amd import build
    if not os.path.exists(build.PROBE_LIB):
        raise RuntimeError(f"{build.PROBE_LIB} is not built (run __graft_entry__.build())")
    lib = C.CDLL(build.PROBE_LIB)
    lib.ems_hbm_probe.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_int,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.ems_hbm_probe_error.restype = C.c_char_p
    gbs, ms = C.c_double(), C.c_double()
    rc = lib.ems_hbm_probe(reads, writes, int(nontemporal), int(bytes_per_array), reps, C.byref(gbs), C.byref(ms))
    if rc != 0:
        raise RuntimeError(lib.ems_hbm_probe_error().decode())
    return gbs.value, ms.value


HBM_PROBE_SWEEP = [  # (reads, writes): read-only, write-only, copy, and the update kernels' mixes
    (1, 0), (4, 0), (0, 1), (0, 2), (1, 1), (2, 1), (7, 3), (4, 3)]


def hbm_probe_sweep(sizes_gib=(1, 2, 4, 8)):
    """Bandwidth of the streaming probe over array sizes and read/write mixes,
    with and without non-temporal hints; one JSON line per case."""
    rows = []
    for gib in sizes_gib:
        for r, w in HBM_PROBE_SWEEP:
            for nt in (0, 1):
                per = int(gib * (1 << 30) / (r + w)) & ~0xFFFF     # total footprint = gib GiB
                gbs, ms = hbm_probe(r, w, nt, per)
                row = {"hbm_probe": {"reads": r, "writes": w, "nontemporal": nt, "total_gib": gib,
                                     "bytes_per_launch": per * (r + w), "ms": round(ms, 4), "GBs": round(gbs, 1)}}
                print(json.dumps(row), flush=True)
                rows.append(row["hbm_probe"])
    return rows


def achievable_bandwidth(total_bytes=3 << 30):
    """Best streaming rate of the probe at the update kernels' read/write mix
    (7 arrays read, 3 written: a half step reads 2.2-2.4 bytes per byte it
    writes) and for a plain copy, with the footprint of the 256^3 fields."""
    best = {}
    for name, (r, w) in (("mix_7r3w", (7, 3)), ("copy", (1, 1))):
        per = int(total_bytes / (r + w)) & ~0xFFFF
        res = [hbm_probe(r, w, nt, per) for nt in (0, 1)]
        best[name] = round(max(g for g, _ in res), 1)
    return best


def _field_digest(eng):
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    h.update(eng.get_volt().tobytes())
    h.update(eng.get_curr().tobytes())
    return h.hexdigest()


def run_batched(count, steps, tune_steps=6, sequential=True):
    """Batched parameter sweep: `count` independent instances of uneven size
    (scenes.batched_sweep, U{48..160}^3 lines, 10 bricks, 6-cell UPML, Gaussian
    soft source, voltage + E-point probes).  Times the instances one after
    another on the default single engine, then all of them in one batch
    (autotuned), and checks that every instance's fields and probe series are
    bit-identical between the two."""
    import numpy as np
    from openems_amd import scenes
    from openems_amd.batch import BatchEngine
    from openems_amd.simulation import Simulation
    scs = scenes.batched_sweep(count=count, nrts=steps)
    cells = sum(int(np.prod([len(l) for l in s["lines"]])) for s in scs)
    seq_s = 0.0
    digests, tds = [], []
    for sc in (scs if sequential else []):
        sim = Simulation(sc)
        sim.eng.sync()
        t0 = time.perf_counter()
        sim.run(steps)
        sim.eng.sync()
        seq_s += time.perf_counter() - t0
        digests.append(_field_digest(sim.eng))
        tds.append({k: v["td"] for k, v in sim.results()["probes"].items()})
        sim.close()
    print(f"# batched leg: sequential {seq_s:.2f} s for {count} instances", file=sys.stderr, flush=True)
    # the same instances one after another on the plain-coefficient engine
    seq_plain_s = 0.0
    for sc in (scs if sequential else []):
        sim = Simulation(sc, layout="plain")
        sim.eng.sync()
        t0 = time.perf_counter()
        sim.run(steps)
        sim.eng.sync()
        seq_plain_s += time.perf_counter() - t0
        sim.close()
    b = BatchEngine(scs)
    tuned_ms = b.autotune(steps=tune_steps)
    print(f"# batched leg: autotuned {b.config()['tile_x']=} {tuned_ms:.3f} ms/step", file=sys.stderr, flush=True)
    cfg = b.config()
    t0 = time.perf_counter()
    b.iterate(steps)
    bat_s = time.perf_counter() - t0
    same = bool(sequential)
    for sim, d, td in zip(b.sims, digests, tds):
        if _field_digest(sim.eng) != d:
            same = False
        got = sim.results()["probes"]
        for k, v in td.items():
            if not np.array_equal(got[k]["td"], v):
                same = False
    out = {"workload": f"batched_sweep: {count} instances, U{{48..160}}^3 lines (total {cells} cells), "
                       f"10 bricks, 6-cell UPML, Gaussian soft source, V + E probes, {steps} steps",
           "instances": count, "cells": cells, "steps": steps,
           "sequential_s": round(seq_s, 3) if sequential else None, "batched_s": round(bat_s, 3),
           "speedup": round(seq_s / bat_s, 3) if sequential else None,
           "sequential_plain_s": round(seq_plain_s, 3) if sequential else None,
           "speedup_vs_plain_engine": round(seq_plain_s / bat_s, 3) if sequential else None,
           "batched_mcells": round(cells * steps / bat_s / 1e6, 1),
           "sequential_mcells": round(cells * steps / seq_s / 1e6, 1) if sequential else None,
           "bit_identical": bool(same), "layout": b.layout,
           "autotuned": {"tile_x": cfg["tile_x"], "grid": cfg["grid"], "work_items": cfg["work_items"],
                         "groups": cfg["groups"], "group_bytes": cfg["group_bytes"], "streams": b.streams(),
                         "zdepths": sorted(set(cfg["zdepth"])), "tuned_ms_per_step": round(tuned_ms, 4)},
           "sequential_engine": "default single engine (compressed coefficients, two half steps), one "
                                "instance after another, setup excluded"}
    b.close()
    return out


def run_adjoint(nsteps=8000, K=256, seed=5):
    """Adjoint gradient workload: 128x128x96 lines at 0.25 mm, 8-cell UPML,
    Gaussian E_z sheet source (fc = 40 GHz), 48x48x8-cell design region
    (eps_r ~ U[1,12], kappa = 0), J = sum over 16 E_z probes of |DFT(30 GHz)|^2,
    8000 steps, checkpoints every 256 steps.  Times one plain forward run
    (default engine, no checkpoints) against the forward with checkpoints plus
    the backward pass, and checks the full gradient along a random +-1
    direction with a centred difference (step 1e-2) of float32 forward runs."""
    import numpy as np
    from openems_amd import scenes
    from openems_amd.adjoint import DesignProblem
    from openems_amd.operator import Operator
    from openems_amd.engine import EngineHIP
    scene, probes, freq = scenes.adjoint_design(nrts=nsteps)
    eps = scene["materials"][0]["grid"]["eps"].copy()
    # plain forward
    op = Operator(scene)
    eng = EngineHIP(op)
    eng.sync()
    t0 = time.perf_counter()
    eng.IterateTS(nsteps)
    eng.sync()
    fwd_s = time.perf_counter() - t0
    eng.close()
    op.close()
    prob = DesignProblem(scene, probes, freq, nsteps, checkpoint_every=K)
    prob._prepare(eps, None)
    import ctypes as C
    from openems_amd._native import check, dp, lib
    J = C.c_double()
    prob._eng.sync()
    t0 = time.perf_counter()
    check(lib().ems_adj_forward(prob._adj, nsteps, C.byref(J)))
    prob._eng.sync()
    fwd_ck_s = time.perf_counter() - t0
    nb = int(np.prod(prob.box_hi - prob.box_lo + 1))
    gvi, gvv = np.zeros(3 * nb), np.zeros(3 * nb)
    t0 = time.perf_counter()
    check(lib().ems_adj_backward(prob._adj, gvi.ctypes.data_as(dp), gvv.ctypes.data_as(dp)))
    bwd_s = time.perf_counter() - t0
    ge, gk = prob._chain(gvi, gvv)
    print(f"# adjoint leg: forward {fwd_s:.3f} s, checkpointed forward {fwd_ck_s:.3f} s, backward {bwd_s:.3f} s",
          file=sys.stderr, flush=True)
    d = np.random.default_rng(seed).choice([-1.0, 1.0], size=eps.shape)
    h = 1e-2
    Jp = prob.forward(eps + h * d)
    Jm = prob.forward(eps - h * d)
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.sum(ge * d))
    prob.close()
    return {"workload": f"adjoint_design: 128x128x96 lines at 0.25 mm, 8-cell UPML, 48x48x8 design cells, "
                        f"16 E_z probes, J = sum |DFT(30 GHz)|^2, {nsteps} steps, checkpoints every {K}",
            "J": J.value, "forward_s": round(fwd_s, 4), "checkpointed_forward_s": round(fwd_ck_s, 4),
            "backward_s": round(bwd_s, 4), "adjoint_over_forward": round((fwd_ck_s + bwd_s) / fwd_s, 3),
            "directional_check": {"direction": "random +-1 per design cell", "step": h,
                                  "adjoint": ad, "central_difference": fd,
                                  "rel_err": abs(ad - fd) / max(abs(fd), 1e-300)},
            "grad_norm_eps": float(np.linalg.norm(ge))}


def run_memory_scale(steps=200, warmup=20):
    """Memory-scale workload: 1024x1024x768 lines (805 M cells), 12 seeded
    dielectric layers, Mur on every face, sinusoidal 15 GHz E_x sheet source,
    16 E-point probes and a z-line voltage probe, steady-state end criterion
    off; default engine (compressed layout).  Throughput over `steps` steps
    after `warmup`, with the update kernels' average times (HIP events)."""
    from openems_amd import scenes
    from openems_amd.simulation import Simulation
    scene = scenes.mur_layers(nrts=steps + warmup)
    scene["steady_state"] = False
    t0 = time.perf_counter()
    sim = Simulation(scene)
    build_s = time.perf_counter() - t0
    eng, op = sim.eng, sim.op
    cells = float(op.num_cells)
    lay, nu, _, _ = eng.layout()
    quad, qz = eng.quad()
    form, fz = eng.fused()
    eng.IterateTS(warmup)
    eng.sync()
    eng.set_profiling(True)
    t0 = time.perf_counter()
    eng.IterateTS(steps)
    eng.sync()
    wall = time.perf_counter() - t0
    _, vms, cms = eng.last_timing()
    eng.set_profiling(False)
    sim.close()
    if form == 3:
        # one sweep per step (volt_ms: the sweep; curr_ms: the Mur hooks and
        # the face-band currents after it)
        kms = vms
        kname = "k_step4t" if op.N[0] // 4 >= 64 else "k_step4"
        per_launch = FUSED_BYTES[lay] * cells
        layout = f"{lay} ({nu} unique), quad column sweep in 64 x 4 quad tiles, z depth {fz}"
        cell_bytes = FUSED_BYTES[lay]
    else:
        kms = max(vms, cms)
        kname = ("k_curr4" if quad else "k_curr") if cms >= vms else ("k_volt4" if quad else "k_volt")
        per_launch = LAYOUT_BYTES[lay] * cells
        layout = f"{lay} ({nu} unique), two half steps, {'four x cells per lane' if quad else 'one cell per lane'}"
        cell_bytes = 2 * LAYOUT_BYTES[lay]
    return {"workload": "mur_layers: 1024x1024x768 lines, 12 seeded layers, Mur x6, sinusoidal E_x sheet, "
                        f"16 E probes + 1 voltage probe, {steps} steps", "cells": int(cells),
            "layout": layout,
            "value": round(cells * steps / wall / 1e6, 1), "unit": "MCells/s", "ms_per_step": round(wall * 1e3 / steps, 3),
            "setup_s": round(build_s, 1),
            "update_kernel": kname if form == 3 else ("k_volt4/k_curr4" if quad else "k_volt/k_curr"),
            "update_kernel_ms": round(kms, 4), "volt_ms": round(vms, 4), "curr_ms": round(cms, 4),
            "bytes_per_launch": int(per_launch), "bytes_per_cell_update": cell_bytes,
            "achieved_GBs": round(per_launch / (kms * 1e-3) / 1e9, 1),
            "frac": round(per_launch / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "dominant_kernel": kname,
            "traffic": traffic_from_profile(kname, layout, pattern="mur_layers_1024x1024x768_traffic*.json")}


def run_plain_cavity(steps=2000, cpu=True, nthreads=1):
    """Plain EC-FDTD workload: 64^3 lines at 1 mm, vacuum, PEC on every face,
    soft E_z Gaussian (f0 = 0, fc = 10 GHz) at the centre, E_z point probe and
    z-line voltage probe, 2000 steps (default engine), with the restated CPU
    engines on the same operator."""
    from openems_amd import scenes
    from openems_amd.simulation import Simulation
    scene = scenes.pec_cavity(64, steps)
    sim = Simulation(scene)
    eng = sim.eng
    cells = float(sim.op.num_cells)
    eng.IterateTS(20)
    eng.sync()
    t0 = time.perf_counter()
    eng.IterateTS(steps)
    eng.sync()
    wall = time.perf_counter() - t0
    quad, _ = eng.quad()
    lay = eng.layout()[0]
    out = {"workload": f"pec_cavity_64: 64^3 lines, 1 mm, vacuum, PEC x6, soft E_z Gaussian fc=10 GHz, "
                       f"E point + voltage line probes, {steps} steps",
           "cells": int(cells), "value": round(cells * steps / wall / 1e6, 1), "unit": "MCells/s",
           "ms_per_step": round(wall * 1e3 / steps, 5),
           "layout": f"{lay}, two half steps, {'four x cells per lane' if quad else 'one cell per lane'}, "
                     f"timestep chunks replayed as HIP graphs",
           "note": "3.1 MB per field: cache-resident and launch-bound, no HBM roofline"}
    if cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(sim.op, nthreads, budget_s=4.0, single_budget_s=4.0)
        except Exception as e:  # the baseline is reported, not required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    sim.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--layout", default="auto", choices=["auto", "plain", "compressed"])
    ap.add_argument("--step", default="auto", choices=["auto", "halves", "lds", "columns", "quads"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--batched-count", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=4000)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched leg without the one-after-another reference runs (for profiling the batch)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-memory-scale", action="store_true",
                    help="skip the 1024x1024x768 Mur workload (its operator build holds about 80 GB of host memory)")
    ap.add_argument("--memory-scale-steps", type=int, default=200)
    ap.add_argument("--no-plain-cavity", action="store_true", help="skip the 64^3 PEC cavity workload")
    ap.add_argument("--no-achievable", action="store_true",
                    help="skip the streaming-probe measurement of the achievable HBM bandwidth")
    ap.add_argument("--adjoint-steps", type=int, default=8000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (for scaling HBM byte counters)")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only run the streaming bandwidth probe sweep (1-8 GiB footprints, read/write mixes)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.hbm_probe:
        hbm_probe_sweep()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine runs on a single MI355X in this build)", file=sys.stderr)
        sys.exit(2)

    from openems_amd import scenes
    from openems_amd.simulation import Simulation

    n = args.size
    scene = scenes.upml_bricks(n, 10000)
    t0 = time.perf_counter()
    sim = Simulation(scene, layout=args.layout,
                     fused=None if args.step == "auto" else args.step)
    setup_s = time.perf_counter() - t0
    eng = sim.eng
    op = sim.op
    cells = float(op.num_cells)
    npml = pml_positions(op)

    eng.IterateTS(args.warmup)
    eng.sync()
    eng.set_profiling(True)
    t0 = time.perf_counter()
    eng.IterateTS(args.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    dev_ms, volt_ms, curr_ms = eng.last_timing()
    eng.set_profiling(False)

    value = cells * args.steps / wall / 1e6
    lay, n_unique, play, n_unique_pml = eng.layout()
    play = play or "plain"
    fused, zdepth = eng.fused()
    quad, qzdepth = eng.quad()
    if fused:
        bytes_per_launch = FUSED_BYTES[lay] * cells + FUSED_PML_BYTES[play] * npml
        step_bytes = bytes_per_launch
    else:
        bytes_per_launch = LAYOUT_BYTES[lay] * cells + PML_BYTES[play] * npml
        step_bytes = 2 * bytes_per_launch
    plain_step_bytes = PLAIN_CELL_STEP_BYTES * (cells + npml)
    if fused:
        kname, kms = ("k_step", "k_step2", "k_step4")[fused - 1], volt_ms
    else:
        kname, kms = ("k_volt", volt_ms) if volt_ms >= curr_ms else ("k_curr", curr_ms)
        if quad:
            kname += "4"
    achieved = bytes_per_launch / (kms * 1e-3) / 1e9
    layout = f"{lay} coefficients ({n_unique} unique) + fused {play} UPML ({n_unique_pml} unique)" \
        if lay == "compressed" or play == "compressed" else "plain fp32 coefficients + fused UPML"
    form = ("two half steps", "single sweep, LDS planes (k_step)", "single sweep, independent columns (k_step2)",
            "single sweep, quad columns (k_step4)")[fused]
    if quad:
        form += ", four x cells per lane"
        zdepth = qzdepth
    layout += f", {form} (z depth {zdepth})"
    traffic = traffic_from_profile(kname, layout, n)
    plain_ceiling = HBM_PEAK_GBS * 1e9 / (plain_step_bytes / cells) / 1e6     # MCells/s
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MCells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"upml_bricks_{n}: {n}^3 lines, 0.5 mm core graded x1.2 over 10 cells, "
                               "40 seeded dielectric bricks, 8-cell UPML on all faces, 50 Ohm lumped port "
                               "with Gaussian fc=20 GHz, port V/I probes, E-field DFT plane at 5/10/15 GHz",
                   "cells": int(cells), "pml_positions": int(npml), "layout": layout,
                   "bytes_per_cell_update": round(step_bytes / cells, 2),
                   "plain_bytes_per_cell_update": round(plain_step_bytes / cells, 2)},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(kms, 5),
                     "volt_ms": round(volt_ms, 5), "curr_ms": round(curr_ms, 5),
                     "plain_ceiling_mcells": round(plain_ceiling, 1),
                     "plain_ceiling_frac": round(value / plain_ceiling, 4)},
    }
    print(f"# setup {setup_s:.1f}s, device time {dev_ms:.1f} ms for {args.steps} steps", file=sys.stderr, flush=True)
    nthreads = cpu_threads()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(op, nthreads, args.cpu_seconds)
        except Exception as e:  # the baseline is reported, not required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
        print(f"# cpu baseline: {out['cpu_baseline']}", file=sys.stderr, flush=True)
    sim.close()
    if not args.no_achievable:
        try:
            ach = achievable_bandwidth()
            out["roofline"]["achievable"] = ach
            out["roofline"]["achievable_unit"] = "GB/s (streaming probe, 16 B per lane, 3 GiB footprint)"
            out["roofline"]["frac_of_achievable"] = round(achieved / ach["mix_7r3w"], 4)
        except Exception as e:  # reported, the headline line still prints
            out["roofline"]["achievable"] = {"error": str(e)}
        print(f"# achievable: {out['roofline'].get('achievable')}", file=sys.stderr, flush=True)
    if not args.no_plain_cavity:
        try:
            out["plain_cavity"] = run_plain_cavity(cpu=not args.no_cpu_baseline, nthreads=nthreads)
        except Exception as e:  # reported, the headline line still prints
            out["plain_cavity"] = {"error": str(e)}
        print(f"# plain cavity: {out['plain_cavity']}", file=sys.stderr, flush=True)
    if not args.no_batched:
        try:
            out["batched"] = run_batched(args.batched_count, args.batched_steps, sequential=not args.no_sequential)
        except Exception as e:  # reported, the headline line still prints
            out["batched"] = {"error": str(e)}
        print(f"# batched: {out['batched']}", file=sys.stderr)
    if not args.no_adjoint:
        try:
            out["adjoint"] = run_adjoint(args.adjoint_steps)
        except Exception as e:  # reported, the headline line still prints
            out["adjoint"] = {"error": str(e)}
        print(f"# adjoint: {out['adjoint']}", file=sys.stderr)
    if not args.no_memory_scale:
        try:
            out["memory_scale"] = run_memory_scale(args.memory_scale_steps)
        except Exception as e:  # reported, the headline line still prints
            out["memory_scale"] = {"error": str(e)}
        print(f"# memory scale: {out['memory_scale']}", file=sys.stderr)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
