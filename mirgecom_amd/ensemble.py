"""Batched ensemble of independent reacting Navier-Stokes cases on one GPU.

The reference runs one driver per case (mirgecom/steppers.py:350-449, RHS
mirgecom/navierstokes.py:350-575 with the chemistry sources of
mirgecom/eos.py:858-875).  Here many cases of uneven mesh size and
polynomial order share every launch: one RHS of all cases is four launches
(Newton temperature, gradients + nodal net flux, divergence, species
sources) over per-case descriptors and work items (``mg_ensemble_prepare`` /
``mg_ensemble_ns_rhs``, include/mgdg.h), each case's states live in one
contiguous buffer of the whole ensemble, and an RK4 step (four RHS and the
FP64 stage updates over that buffer) is captured once in a HIP graph.

Cases: periodic boxes [0, 2 pi]^3 of n^3 x 6 tets, order p, the uiuc_7sp
mixture with constant transport set by a Reynolds number, initial fields of
the Taylor-Green kind (velocity and pressure of the Taylor-Green vortex
shifted by a seeded phase) with the reacting box's seeded temperature and
mixture-fraction fields; mixed-precision contractions, FP64 Newton
temperature and RK accumulation, and the learned chemistry surrogate (with
its gate) or the exact rates.
"""

from __future__ import annotations

import json
import math
import os
import time
from dataclasses import dataclass

import numpy as np
import torch

from mirgecom_amd import _lib
from mirgecom_amd import mechanisms as mech

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                           "ensemble_tuning.json")
NFIELDS = 12
RHO0 = 0.22          # kg/m^3, ethylene-air near 1600 K and 1 atm
V0 = 10.0            # m/s, Taylor-Green velocity scale
P0 = 101325.0


@dataclass
class CaseSpec:
    n: int
    order: int
    reynolds: float
    seed: int


def random_specs(ncases=48, seed=0, n_range=(8, 28), orders=(1, 2, 3, 4),
                 re_range=(500.0, 5000.0)):
    """Seeded case list: n uniform in n_range, p uniform in orders, Re
    uniform in re_range."""
    rng = np.random.default_rng(seed)
    return [CaseSpec(int(rng.integers(n_range[0], n_range[1] + 1)),
                     int(rng.choice(orders)), float(rng.uniform(*re_range)),
                     int(seed * 1000 + i))
            for i in range(ncases)]


def case_gas_model(reynolds, prandtl=0.71, schmidt=1.0):
    """uiuc_7sp mixture with constant mu = rho0 V0 L / Re (L = 1),
    kappa = mu cp / Pr and d_k = mu / (rho0 Sc)."""
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.gas_model import GasModel
    from mirgecom_amd.transport import SimpleTransport
    mu = RHO0 * V0 / reynolds
    cp = 1400.0
    tr = SimpleTransport(viscosity=mu, thermal_conductivity=mu * cp / prandtl,
                         species_diffusivity=np.full(7, mu / (RHO0 * schmidt)))
    return GasModel(eos=PyrometheusMixture(temperature_niter=5), transport=tr)


def case_initial_state(dcoll, seed):
    """Taylor-Green velocity and pressure (phase-shifted by the seed) on the
    reacting box's seeded temperature / mixture-fraction fields."""
    from mirgecom_amd.eos import PyrometheusMixture
    from mirgecom_amd.fluid import make_conserved
    from mirgecom_amd.workloads import _fourier_field, stoichiometric_mixture_fraction
    x = dcoll.nodes()
    rng = np.random.default_rng(seed)
    sh = rng.uniform(0, 2 * math.pi, 3)
    xs = [x[i] + float(sh[i]) for i in range(3)]
    u = V0 * torch.sin(xs[0]) * torch.cos(xs[1]) * torch.cos(xs[2])
    v = -V0 * torch.cos(xs[0]) * torch.sin(xs[1]) * torch.cos(xs[2])
    w = torch.zeros_like(u)
    p = P0 + RHO0 * V0**2 / 16 * (torch.cos(2 * xs[0]) + torch.cos(2 * xs[1])) * (
        torch.cos(2 * xs[2]) + 2)
    t = 1600.0 + 300.0 * _fourier_field(x, rng, 6)
    zst = stoichiometric_mixture_fraction()
    z = zst * (1.0 + 0.2 * _fourier_field(x, rng, 4))
    wts = mech.molecular_weights()
    sp = {s: i for i, s in enumerate(mech.SPECIES)}
    yo2 = wts[sp["O2"]] / (wts[sp["O2"]] + 3.76 * wts[sp["N2"]])
    y = torch.zeros((7,) + tuple(t.shape), dtype=t.dtype, device=t.device)
    y[sp["C2H4"]] = z
    y[sp["O2"]] = (1 - z) * yo2
    y[sp["N2"]] = (1 - z) * (1 - yo2)
    eos = PyrometheusMixture()
    rho = eos.get_density(p, t, y)
    energy = rho * eos.get_internal_energy(t, y) + 0.5 * rho * (u * u + v * v + w * w)
    cv = make_conserved(3, mass=rho, energy=energy,
                        momentum=torch.stack([rho * u, rho * v, rho * w]),
                        species_mass=rho * y)
    return cv, t


class ReactingEnsemble:
    """The cases of ``specs`` with one contiguous state buffer, the
    ensemble plans of the three RK4 stage inputs and an optional chemistry
    surrogate (``mirgecom_amd.surrogate.ChemistrySurrogate``; None: exact
    rates)."""

    def __init__(self, specs, surrogate=None, dt=1e-8, tuning=None, device="cuda"):
        from mirgecom_amd.device import get_device_discretization
        from mirgecom_amd.discretization import (
            create_discretization_collection, generate_box_mesh)
        self.specs = list(specs)
        self.dt = float(dt)
        self.surrogate = surrogate
        self.device = torch.device(device)
        self.dcolls, self.gas_models, self.devs = [], [], []
        sizes = []
        for sp in self.specs:
            mesh = generate_box_mesh(3, 0.0, 2 * math.pi, (sp.n,) * 3,
                                     periodic=(True,) * 3)
            dcoll = create_discretization_collection(mesh, sp.order)
            self.dcolls.append(dcoll)
            self.gas_models.append(case_gas_model(sp.reynolds))
            self.devs.append(get_device_discretization(dcoll, self.device))
            sizes.append((mesh.nelements, dcoll.refel.nnodes))
        self.sizes = sizes
        nodes = [k * npe for k, npe in sizes]
        self.nodes = nodes
        self.node_off = np.concatenate([[0], np.cumsum(nodes)])
        tot = int(self.node_off[-1])
        self.total_nodes = tot
        f = NFIELDS

        def buf(width, dtype=torch.float64):
            return torch.zeros(width * tot, dtype=dtype, device=self.device)

        self.y, self.k, self.tmp, self.tmp2, self.acc = (buf(f) for _ in range(5))
        self.temperature = buf(1)
        self.fnet = buf(3 * f, torch.float32)
        self.grad32 = buf(3 * (f + 1), torch.float32)
        self.fallback = torch.zeros(1, dtype=torch.int32, device=self.device)
        for i, sp in enumerate(self.specs):
            cv, t = case_initial_state(self.dcolls[i], sp.seed)
            self.case_view(self.y, i).copy_(cv.q)
            self.case_t(i).copy_(t)
        self.plans = {}
        self.tuning = tuning if tuning is not None else load_tuning()
        self._graph = None
        self.prepare()

    # {{{ buffers

    def case_view(self, buf, i, width=NFIELDS):
        """Case ``i`` of a contiguous ensemble buffer as ``(width, K, Np)``."""
        k, npe = self.sizes[i]
        o = int(self.node_off[i]) * width
        return buf[o:o + width * k * npe].view(width, k, npe)

    def case_t(self, i):
        k, npe = self.sizes[i]
        o = int(self.node_off[i])
        return self.temperature[o:o + k * npe].view(k, npe)

    def case_fnet(self, i):
        return self.case_view(self.fnet, i, 3 * NFIELDS)

    def case_grad(self, i):
        return self.case_view(self.grad32, i, 3 * (NFIELDS + 1))

    # }}}

    # {{{ plans

    def _cases(self, src):
        arr = (_lib.MgEnsCase * len(self.specs))()
        for i in range(len(self.specs)):
            c = arr[i]
            c.mesh = self.devs[i].mesh
            c.gas = self.gas_models[i].device_params(7, with_transport=True)
            c.q = self.case_view(src, i).data_ptr()
            c.t_seed = self.case_t(i).data_ptr()
            c.temperature = self.case_t(i).data_ptr()
            c.fnet = self.case_fnet(i).data_ptr()
            c.grad = self.case_grad(i).data_ptr()
            c.rhs = self.case_view(self.k, i).data_ptr()
        return arr

    def prepare(self, tuning=None):
        """(Re)build the plans of the three stage inputs."""
        if tuning is not None:
            self.tuning = tuning
        _lib.ensure_mechanism()
        lib = _lib.load()
        tun = _lib.MgEnsTuning()
        tun.fused = int(self.tuning.get("fused", 0))
        tun.persistent = int(self.tuning.get("persistent", 0))
        wg = self.tuning.get("wg_per_cu", [4] * 5)
        for p in range(5):
            tun.wg_per_cu[p] = int(wg[p])
        s = self.surrogate
        for name, src in (("y", self.y), ("tmp", self.tmp), ("tmp2", self.tmp2)):
            cases = self._cases(src)
            nbytes = lib.mg_ensemble_plan_bytes(len(cases), cases)
            if nbytes < 0:
                _lib.check(_lib.load().mg_ensemble_prepare(
                    len(cases), cases, 5, self.dt, None, None, None, None, tun,
                    None, 0, None, None))
            old = self.plans.get(name)
            dev = old[1] if old is not None and old[1].numel() >= nbytes else \
                torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            plan = _lib.MgEnsPlan()
            _lib.call("mg_ensemble_prepare", len(cases), cases, 5, self.dt,
                      _lib.ptr(s.params) if s else None,
                      _lib.ptr(s.wbf16) if s else None,
                      s.norm if s else None, _lib.ptr(self.fallback), tun,
                      _lib.ptr(dev), dev.numel(), plan, _lib.stream_handle())
            self.plans[name] = (plan, dev, cases)
        self._graph = None

    # }}}

    # {{{ right-hand sides and steps

    def rhs(self, which="y"):
        """RHS of every case for the stage input ``which`` into ``self.k``."""
        _lib.call("mg_ensemble_ns_rhs", self.plans[which][0], _lib.stream_handle())

    def rhs_sequential(self, which="y"):
        """The same RHS case by case through the single-case entry points
        (Newton, mg_ns_flux_mixed, mg_ns_div_mixed, sources)."""
        src = {"y": self.y, "tmp": self.tmp, "tmp2": self.tmp2}[which]
        st = _lib.stream_handle()
        for i in range(len(self.specs)):
            q = self.case_view(src, i)
            t = self.case_t(i)
            fn = self.case_fnet(i)
            r = self.case_view(self.k, i)
            gas = self.gas_models[i].device_params(7, with_transport=True)
            mesh = self.devs[i].mesh
            n = q.shape[1] * q.shape[2]
            _lib.call("mg_mixture_temperature", gas, 3, n, _lib.ptr(q), _lib.ptr(t), 5,
                      _lib.ptr(t), st)
            _lib.call("mg_ns_flux_mixed", mesh, gas, _lib.ptr(q), _lib.ptr(t), None, None,
                      None, 1, _lib.ptr(fn), None, _lib.ptr(self.case_grad(i)), st)
            _lib.call("mg_ns_div_mixed", mesh, gas, _lib.ptr(q), _lib.ptr(t), _lib.ptr(fn),
                      None, None, None, 0, 1, _lib.ptr(r), st)
            if self.surrogate is not None:
                sg = self.surrogate
                _lib.call("mg_chem_surrogate_source", gas, 3, q.shape[2], n, _lib.ptr(q),
                          _lib.ptr(t), self.dt, _lib.ptr(sg.params), _lib.ptr(sg.wbf16),
                          sg.norm, 1, _lib.ptr(r), _lib.ptr(self.fallback), st)
            else:
                _lib.call("mg_species_source_add", gas, 3, n, _lib.ptr(q), _lib.ptr(t),
                          _lib.ptr(r), st)

    def _axpy(self, y, k, a, out, acc=None, w=0.0):
        _lib.call("mg_rk_axpy", y.numel(), _lib.ptr(y), _lib.ptr(k), float(a),
                  _lib.ptr(out), _lib.ptr(acc), float(w), _lib.stream_handle())

    def step_eager(self, sequential=False):
        """One classical RK4 step of every case (FP64 stage updates and
        accumulation over the whole ensemble buffer)."""
        f = self.rhs_sequential if sequential else self.rhs
        dt = self.dt
        self.acc.copy_(self.y)
        f("y")
        self._axpy(self.y, self.k, dt / 2, self.tmp, self.acc, dt / 6)
        f("tmp")
        self._axpy(self.y, self.k, dt / 2, self.tmp2, self.acc, dt / 3)
        f("tmp2")
        self._axpy(self.y, self.k, dt, self.tmp, self.acc, dt / 3)
        f("tmp")
        self._axpy(self.acc, self.k, dt / 6, self.y)

    def capture(self, warmup=1):
        """Capture one RK4 step in a HIP graph (state in ``self.y``)."""
        saved = self.y.clone()
        tsaved = self.temperature.clone()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self.step_eager()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.y.copy_(saved)
        self.temperature.copy_(tsaved)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.step_eager()
        self.y.copy_(saved)
        self.temperature.copy_(tsaved)
        torch.cuda.synchronize()
        self._graph = g
        return g

    def step(self):
        if self._graph is None:
            self.capture()
        self._graph.replay()

    # }}}

    @property
    def dofs(self):
        return self.total_nodes * NFIELDS


def load_tuning():
    """The committed launch tuning (fused / persistent / workgroups per CU),
    or the untuned default."""
    if os.path.exists(TUNING_PATH):
        with open(TUNING_PATH) as f:
            return json.load(f)["best"]
    return {"fused": 0, "persistent": 0, "wg_per_cu": [4, 4, 4, 4, 4]}


def _time_rhs(ens, reps):
    ens.rhs("y")
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        ens.rhs("y")
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def autotune(ens, reps=10, write=None):
    """Time one ensemble RHS under every launch shape: one launch per phase
    for all orders or one per order; one workgroup per item or persistent
    workgroups (1, 2, 4 or 8 per CU, chosen per order by coordinate
    descent).  Applies the fastest to ``ens`` and returns the table (also
    written to ``write`` as JSON when given)."""
    results = []

    def trial(t):
        ens.prepare(t)
        ms = _time_rhs(ens, reps)
        results.append({"tuning": t, "ms": ms})
        return ms

    base = {"fused": 1, "persistent": 0, "wg_per_cu": [4] * 5}
    best_t, best = base, trial(base)
    t = {"fused": 0, "persistent": 0, "wg_per_cu": [4] * 5}
    ms = trial(t)
    if ms < best:
        best_t, best = t, ms
    for fused in (1, 0):
        wg = [4] * 5
        cur = None
        orders = [0] if fused else [4, 3, 2, 1]
        for p in orders:
            bp, bms = wg[p], None
            for cand in (1, 2, 4, 8):
                w2 = list(wg)
                w2[p] = cand
                t = {"fused": fused, "persistent": 1, "wg_per_cu": w2}
                ms = trial(t)
                if bms is None or ms < bms:
                    bp, bms = cand, ms
            wg[p] = bp
            cur = bms
        if cur is not None and cur < best:
            best_t, best = {"fused": fused, "persistent": 1, "wg_per_cu": wg}, cur
    ens.prepare(best_t)
    table = {"best": best_t, "best_ms": best,
             "key": {"dim": 3, "fields": NFIELDS, "orders": sorted(
                 {s.order for s in ens.specs})},
             "trials": results}
    if write:
        with open(write, "w") as f:
            json.dump(table, f, indent=1)
    return table


def measure(ens, steps=5, sequential=True):
    """Wall time per RK4 step of the shared launches, graph-replayed and
    eager (the faster of the two is the batched figure, both are listed),
    and (optionally) the same step with sequential per-case launches;
    DOF-RHS/s of each."""
    ens.capture()
    ens.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.step()
    torch.cuda.synchronize()
    ms_g = (time.perf_counter() - t0) / steps * 1e3
    ens.step_eager()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.step_eager()
    torch.cuda.synchronize()
    ms_e = (time.perf_counter() - t0) / steps * 1e3
    ms_b = min(ms_g, ms_e)
    out = {"batched_ms_per_step": ms_b,
           "batched_graph_ms_per_step": ms_g,
           "batched_eager_ms_per_step": ms_e,
           "batched_dof_rhs_s": ens.dofs * 4 / (ms_b * 1e-3)}
    if sequential:
        ens.step_eager(sequential=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ens.step_eager(sequential=True)
        torch.cuda.synchronize()
        ms_s = (time.perf_counter() - t0) / steps * 1e3
        out.update({"sequential_ms_per_step": ms_s,
                    "sequential_dof_rhs_s": ens.dofs * 4 / (ms_s * 1e-3),
                    "speedup": ms_s / ms_b})
    return out


__all__ = ["CaseSpec", "ReactingEnsemble", "random_specs", "autotune", "measure",
           "load_tuning", "case_gas_model", "case_initial_state"]
