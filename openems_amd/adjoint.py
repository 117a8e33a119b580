"""Adjoint gradients of a frequency-domain probe objective with respect to
the per-cell eps_r / kappa of a design region (include/ems_hip.h,
``ems_adj_*``), and a ``torch.autograd.Function`` around it.

The objective is J = sum_p |F_p|^2 with
F_p = sum_{n=1..N} (V_p(n) / delta_p) exp(-2 pi i f n dT) dT, the running DFT
at frequency f of the field component E = V / delta at each objective probe
after every step n.  The design region is a material grid of the scene
(per-cell eps_r and kappa).  ``ems_adj_backward`` returns dJ/dvi and dJ/dvv on
the edges around the region; the chain to eps_r and kappa goes through the
operator formulas
    vi = dT / (C + dT G / 2),  vv = (2C - dT G) / (2C + dT G)
(Operator::Calc_ECOperatorPos, FDTD/operator.cpp:956-984) and the
quarter-cell averages C = sum_q dC_q eps_q, G = sum_q dC_q kappa_q / eps0
(Operator::AverageMatQuarterCell + Calc_ECPos, FDTD/operator.cpp:1189-1254,
1447-1510), whose weights dC_q come from the operator builder.
"""
from __future__ import annotations

import copy
import ctypes as C

import numpy as np

from ._native import check, check_op, dp, lib, ptr, up
from .engine import EngineHIP
from .operator import Operator

EPS0 = 8.85418781762e-12


def design_grid(scene):
    """The scene's material-grid entry (the design region)."""
    for m in scene.get("materials", []):
        if "grid" in m:
            return m
    raise ValueError("the scene has no material grid")


class DesignProblem:
    """J(eps_r, kappa) of a scene whose design region is a material grid.

    probes: list of (component, (x, y, z) mesh indices); freq in Hz;
    nsteps timesteps; checkpoint_every steps between saved states."""

    def __init__(self, scene, probes, freq, nsteps, checkpoint_every=256, layout="compressed", device=0):
        self.scene = scene
        self.probes = [(int(c), tuple(int(v) for v in p)) for c, p in probes]
        self.freq = float(freq)
        self.nsteps = int(nsteps)
        self.K = int(checkpoint_every)
        self.layout = layout
        self.device = device
        g = design_grid(scene)["grid"]
        self.cell0 = np.asarray(g["cell0"], np.int64)
        self.ncells = np.asarray(g["ncells"], np.int64)
        self.shape = tuple(int(v) for v in self.ncells[::-1])   # [nz][ny][nx]
        # sensitivity box: edges whose quarter cells can lie in the region
        self.box_lo = self.cell0.copy()
        self.box_hi = self.cell0 + self.ncells
        self._weights = None
        self._op = self._eng = self._adj = None

    # -- scene with given design values ----------------------------------------
    def scene_with(self, eps, kappa=None):
        s = copy.copy(self.scene)
        mats = []
        for m in self.scene["materials"]:
            if "grid" in m:
                m = dict(m)
                g = dict(m["grid"])
                g["eps"] = np.asarray(eps, np.float64).reshape(self.shape)
                g["kappa"] = np.zeros(self.shape) if kappa is None else np.asarray(kappa, np.float64).reshape(self.shape)
                m["grid"] = g
            mats.append(m)
        s["materials"] = mats
        return s

    def _prepare(self, eps, kappa):
        self.close()
        self._op = Operator(self.scene_with(eps, kappa))
        self._eng = EngineHIP(self._op, device=self.device, layout=self.layout, fused=False)
        L = lib()
        np_ = len(self.probes)
        comp = np.array([c for c, _ in self.probes], np.uint32)
        pos = np.array([[p[a] for _, p in self.probes] for a in range(3)], np.uint32)
        div = np.array([self._op.edge_length(c, p[c]) for c, p in self.probes], np.float64)
        lo = (C.c_uint * 3)(*[int(v) for v in self.box_lo])
        hi = (C.c_uint * 3)(*[int(v) for v in self.box_hi])
        a = C.c_void_p()
        self._keep = [comp, pos, div]
        check(L.ems_adj_create(C.byref(a), self._eng.h, np_, comp.ctypes.data_as(up), pos.ctypes.data_as(up),
                               div.ctypes.data_as(dp), self.freq, self._op.dT, lo, hi, self.K))
        self._adj = a

    # -- forward / backward -----------------------------------------------------
    def forward(self, eps, kappa=None):
        """Objective J for the design values (runs the N steps with checkpoints)."""
        self._prepare(eps, kappa)
        J = C.c_double()
        check(lib().ems_adj_forward(self._adj, self.nsteps, C.byref(J)))
        self._eng.sync()
        return J.value

    def backward(self):
        """(dJ/d eps_r, dJ/d kappa) on the design grid after forward()."""
        nb = int(np.prod(self.box_hi - self.box_lo + 1))
        gvi = np.zeros(3 * nb)
        gvv = np.zeros(3 * nb)
        check(lib().ems_adj_backward(self._adj, gvi.ctypes.data_as(dp), gvv.ctypes.data_as(dp)))
        return self._chain(gvi, gvv)

    def value_and_grad(self, eps, kappa=None):
        J = self.forward(eps, kappa)
        ge, gk = self.backward()
        return J, ge, gk

    # -- chain rule to eps_r / kappa ---------------------------------------------
    def _edge_weights(self):
        """For every box edge: the four quarter cells (flat design index or
        -1) and dC/d eps_r of each (geometry only, computed once)."""
        if self._weights is not None:
            return self._weights
        L = lib()
        lo, hi = self.box_lo, self.box_hi
        nbx, nby, nbz = (hi - lo + 1)
        nb = nbx * nby * nbz
        cells = np.full((3, nb, 4), -1, np.int64)
        dC = np.zeros((3, nb, 4))
        cbuf = (C.c_uint * 12)()
        wbuf = (C.c_double * 4)()
        pos = (C.c_uint * 3)()
        nx, ny, nz = self.ncells
        for n in range(3):
            k = 0
            for z in range(lo[2], hi[2] + 1):
                for y in range(lo[1], hi[1] + 1):
                    for x in range(lo[0], hi[0] + 1):
                        pos[0], pos[1], pos[2] = x, y, z
                        check_op(L.emsop_eps_weights(self._op.h, n, pos, cbuf, wbuf))
                        for q in range(4):
                            c = [cbuf[3 * q + a] for a in range(3)]
                            r = [c[a] - self.cell0[a] for a in range(3)]
                            if all(0 <= r[a] < self.ncells[a] for a in range(3)) and 0xFFFFFFFF not in c:
                                cells[n, k, q] = r[0] + nx * (r[1] + ny * r[2])
                                dC[n, k, q] = wbuf[q]
                        k += 1
        self._weights = (cells, dC)
        return self._weights

    def _chain(self, gvi, gvv):
        cells, dC = self._edge_weights()
        lo, hi = self.box_lo, self.box_hi
        nb = cells.shape[1]
        dT = self._op.dT
        Cc = self._op.ec("C")
        Gc = self._op.ec("G")
        vi = self._op.coeff("vi")
        sl = (slice(lo[0], hi[0] + 1), slice(lo[1], hi[1] + 1), slice(lo[2], hi[2] + 1))
        ncell = int(np.prod(self.ncells))
        ge = np.zeros(ncell)
        gk = np.zeros(ncell)
        for n in range(3):
            # box arrays in x-fastest order, like the engine's box numbering
            Cn = Cc[n][sl].transpose(2, 1, 0).reshape(-1).astype(np.float64)
            Gn = Gc[n][sl].transpose(2, 1, 0).reshape(-1).astype(np.float64)
            live = vi[n][sl].transpose(2, 1, 0).reshape(-1) != 0   # coefficients set from C, G (not zeroed)
            den1 = Cn + dT * Gn / 2
            den2 = 2 * Cn + dT * Gn
            with np.errstate(divide="ignore", invalid="ignore"):
                dvi_dC = np.where(live, -dT / den1 ** 2, 0.0)
                dvi_dG = np.where(live, -(dT * dT / 2) / den1 ** 2, 0.0)
                dvv_dC = np.where(live, 4 * dT * Gn / den2 ** 2, 0.0)
                dvv_dG = np.where(live, -4 * Cn * dT / den2 ** 2, 0.0)
            a = gvi[n * nb:(n + 1) * nb]
            b = gvv[n * nb:(n + 1) * nb]
            dJdC = a * dvi_dC + b * dvv_dC
            dJdG = a * dvi_dG + b * dvv_dG
            for q in range(4):
                m = cells[n, :, q] >= 0
                np.add.at(ge, cells[n, m, q], dJdC[m] * dC[n, m, q])
                np.add.at(gk, cells[n, m, q], dJdG[m] * dC[n, m, q] / EPS0)
        return ge.reshape(self.shape), gk.reshape(self.shape)

    def close(self):
        if self._adj is not None:
            lib().ems_adj_destroy(self._adj)
            self._adj = None
        if self._eng is not None:
            self._eng.close()
            self._eng = None
        if self._op is not None:
            self._op.close()
            self._op = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def objective(eps, kappa, problem):
    """Differentiable J(eps_r, kappa) (torch tensors shaped like the design grid)."""
    import torch

    class _Objective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, e, k):
            J = problem.forward(e.detach().cpu().double().numpy(), k.detach().cpu().double().numpy())
            ctx.meta = (e.dtype, e.device, k.dtype, k.device)
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            ge, gk = problem.backward()
            ed, edev, kd, kdev = ctx.meta
            g = gJ.detach().cpu().double()
            return ((g * torch.from_numpy(ge)).to(dtype=ed, device=edev),
                    (g * torch.from_numpy(gk)).to(dtype=kd, device=kdev))

    return _Objective.apply(eps, kappa)
