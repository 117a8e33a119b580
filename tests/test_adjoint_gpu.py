"""Adjoint gradients (include/ems_hip.h ems_adj_*, openems_amd.adjoint)
against centred finite differences of the oracle's float64 engine on the
same operator: the reference gradient in double precision on the CPU."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.adjoint import DesignProblem, objective
from oracle.operator_ref import RefOperator
from oracle.oracle import RefEngine

pytestmark = pytest.mark.gpu


def _small():
    return scenes.adjoint_design(30, 28, 26, pml=6, design=(6, 5, 3), nrts=400, fc=60e9, freq=40e9)


def oracle_J(scene, probes, freq, nsteps):
    """J = sum_p |sum_n V_p(n)/delta exp(-2 pi i f n dT) dT|^2 with the oracle's
    float64 engine (float32 operator as built by the reference)."""
    ref = RefOperator(scene).build()
    eng = RefEngine(ref, "basic_f64")
    V, _ = eng.raw_views()
    d = [np.diff(np.asarray(l) * scene["delta_unit"]) for l in scene["lines"]]
    F = np.zeros(len(probes), np.complex128)
    for n in range(1, nsteps + 1):
        eng.iterate(1)
        w = np.exp(-2j * np.pi * freq * n * ref.dT) * ref.dT
        for k, (c, p) in enumerate(probes):
            F[k] += V[c][p] / d[c][p[c]] * w
    eng.close()
    return float(np.sum(np.abs(F) ** 2))


def test_objective_matches_oracle_float64(gpu):
    scene, probes, freq = _small()
    print("case adjoint small: objective vs oracle float64", flush=True)
    prob = DesignProblem(scene, probes, freq, 400, checkpoint_every=64)
    eps = scene["materials"][0]["grid"]["eps"]
    J = prob.forward(eps)
    Jref = oracle_J(scene, probes, freq, 400)
    assert Jref > 0
    assert J == pytest.approx(Jref, rel=1e-4)
    prob.close()


def test_gradient_against_central_differences(gpu):
    """dJ/d eps_r and dJ/d kappa on random design cells vs centred FD
    (steps 1e-2 in eps_r and in kappa [S/m], around kappa = 0.02 S/m so that
    vv depends on the design too) of the float64 oracle; agreement within
    1e-3 (eps_r) and 2e-3 (kappa) of the largest gradient entry."""
    scene, probes, freq = _small()
    print("case adjoint small: gradient vs FD float64", flush=True)
    prob = DesignProblem(scene, probes, freq, 400, checkpoint_every=64)
    eps = scene["materials"][0]["grid"]["eps"].copy()
    kap = np.full_like(eps, 0.02)
    J, ge, gk = prob.value_and_grad(eps, kap)
    assert np.all(np.isfinite(ge)) and np.abs(ge).max() > 0
    rng = np.random.default_rng(9)
    flat = rng.choice(eps.size, size=6, replace=False)
    h = 1e-2
    scale = np.abs(ge).max()
    for f in flat:
        idx = np.unravel_index(f, eps.shape)
        ep, em = eps.copy(), eps.copy()
        ep[idx] += h
        em[idx] -= h
        Jp = oracle_J(prob.scene_with(ep, kap), probes, freq, 400)
        Jm = oracle_J(prob.scene_with(em, kap), probes, freq, 400)
        fd = (Jp - Jm) / (2 * h)
        print("cell", idx, "adjoint", ge[idx], "fd", fd, flush=True)
        assert abs(ge[idx] - fd) <= 1e-3 * scale
    # kappa: conductivity steps of 1e-2 S/m
    hk = 1e-2
    scale_k = np.abs(gk).max()
    for f in flat[:2]:
        idx = np.unravel_index(f, eps.shape)
        kp, km = kap.copy(), kap.copy()
        kp[idx] += hk
        km[idx] -= hk
        Jp = oracle_J(prob.scene_with(eps, kp), probes, freq, 400)
        Jm = oracle_J(prob.scene_with(eps, km), probes, freq, 400)
        fd = (Jp - Jm) / (2 * hk)
        print("cell", idx, "adjoint kappa", gk[idx], "fd", fd, flush=True)
        assert abs(gk[idx] - fd) <= 2e-3 * scale_k
    prob.close()


def test_checkpoint_interval_does_not_change_gradient(gpu):
    scene, probes, freq = _small()
    eps = scene["materials"][0]["grid"]["eps"]
    out = []
    for K in (37, 400):
        prob = DesignProblem(scene, probes, freq, 400, checkpoint_every=K)
        out.append(prob.value_and_grad(eps))
        prob.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def test_torch_autograd_function(gpu):
    import torch
    scene, probes, freq = _small()
    prob = DesignProblem(scene, probes, freq, 400, checkpoint_every=64)
    eps = scene["materials"][0]["grid"]["eps"]
    e = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    k = torch.zeros_like(e, requires_grad=True)
    J = objective(e, k, prob)
    J.backward()
    J2, ge, gk = prob.value_and_grad(eps, np.zeros_like(eps))
    assert J.item() == J2
    assert np.array_equal(e.grad.numpy(), ge) and np.array_equal(k.grad.numpy(), gk)
    prob.close()


def test_gradient_32_cells_against_fp64_central_differences(gpu):
    """The finite-difference check of the adjoint workload at the largest
    size the CPU reference affords in a test: 64 x 64 x 48 lines at 0.25 mm,
    8-cell UPML, 24 x 24 x 4 design cells with eps_r ~ U[1, 12], 2000 steps.
    dJ/d eps_r from the GPU adjoint (float32 fields, checkpoints every 256
    steps) against centred differences with step 1e-3 of the double-precision
    reference forward (oracle.adjoint_ref) on 32 random design cells:
    every cell within 1e-3 of the largest difference quotient, and the 32-cell
    vectors within 1e-3 relative (2-norm)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from oracle.adjoint_ref import objective_f64_job
    scene, probes, freq = scenes.adjoint_design(64, 64, 48, pml=8, design=(24, 24, 4), nrts=2000)
    print("case adjoint 64x64x48: 32 cells vs FP64 centred differences (step 1e-3)", flush=True)
    prob = DesignProblem(scene, probes, freq, 2000, checkpoint_every=256)
    eps = scene["materials"][0]["grid"]["eps"].copy()
    J, ge, _ = prob.value_and_grad(eps)
    assert J > 0 and np.all(np.isfinite(ge))
    rng = np.random.default_rng(32)
    flat = rng.choice(eps.size, size=32, replace=False)
    h = 1e-3
    lo = tuple(int(v) for v in prob.box_lo - 1)
    hi = tuple(int(v) for v in prob.box_hi + 1)
    jobs = []
    for f in flat:
        idx = np.unravel_index(f, eps.shape)
        for s in (+1, -1):
            e = eps.copy()
            e[idx] += s * h
            jobs.append((prob.scene_with(e), probes, freq, 2000, lo, hi))
    prob.close()
    with ProcessPoolExecutor(max_workers=16, mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(objective_f64_job, jobs))
    dts = {r[1] for r in res}
    assert len(dts) == 1, "the design must not change the timestep"
    fd = np.array([(res[2 * k][0] - res[2 * k + 1][0]) / (2 * h) for k in range(len(flat))])
    ad = np.array([ge[np.unravel_index(f, eps.shape)] for f in flat])
    scale = np.abs(fd).max()
    for f, a, d in zip(flat, ad, fd):
        print("cell", np.unravel_index(f, eps.shape), "adjoint", a, "fd64", d, flush=True)
    assert scale > 0
    assert np.max(np.abs(ad - fd)) <= 1e-3 * scale
    assert np.linalg.norm(ad - fd) <= 1e-3 * np.linalg.norm(fd)


def test_full_size_directional_derivative(gpu):
    """The adjoint workload at full size (128 x 128 x 96 lines, 48 x 48 x 8
    design cells, 8000 steps, checkpoints every 256): the gradient along a
    random +-1 direction against a centred difference (step 1e-2) of two
    float32 forward runs, within 1e-3 relative."""
    scene, probes, freq = scenes.adjoint_design(nrts=8000)
    print("case adjoint full size directional derivative", flush=True)
    prob = DesignProblem(scene, probes, freq, 8000, checkpoint_every=256)
    eps = scene["materials"][0]["grid"]["eps"].copy()
    J, ge, _ = prob.value_and_grad(eps)
    d = np.random.default_rng(5).choice([-1.0, 1.0], size=eps.shape)
    h = 1e-2
    Jp = prob.forward(eps + h * d)
    Jm = prob.forward(eps - h * d)
    prob.close()
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.sum(ge * d))
    print("J", J, "adjoint", ad, "fd", fd, flush=True)
    assert J > 0 and fd != 0
    assert abs(ad - fd) <= 1e-3 * abs(fd)
