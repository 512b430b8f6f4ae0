"""Benchmark: DG right-hand-side throughput on one MI355X.

Workload (default): 3-D compressible Navier-Stokes Taylor-Green vortex,
periodic cube [0, 2 pi]^3 of 24^3 cells x 6 tets = 82,944 tets, p=3
(20 nodes/element, 1,658,880 nodes x 5 fields), Re=1600, Ma=0.1, Pr=0.71,
constant viscosity, ideal gas gamma=1.4, Rusanov inviscid flux and central
viscous flux.  One "step" is one full RHS evaluation (``ns_operator``: the
gradient kernel and the divergence kernel) with the state resident in HBM.

Prints ONE JSON line with the metric (DOF-RHS/s = nodes x fields / RHS wall
time), the roofline fraction of the dominant kernel (algorithmic HBM bytes
per launch / its average event-timed duration / 8 TB/s) and the CPU baseline
(the FP64 NumPy oracle on a bounded sample of the same workload, host cores).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=24, help="cells per axis")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=12,
                    help="cells per axis of the CPU-baseline sample")
    ap.add_argument("--state-steps", type=int, default=200,
                    help="RK4 steps of the mixed-vs-FP64 state check")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the HIP-graph replay timing")
    ap.add_argument("--reacting-n", type=int, default=32,
                    help="cells per axis of the reacting-box line (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (3x), the "
                         "reference kernel for HBM-traffic counter scaling")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the chemistry-surrogate leg")
    ap.add_argument("--sg-reactors", type=int, default=4096,
                    help="autoignition reactors of the surrogate training set")
    ap.add_argument("--sg-samples", type=int, default=1024,
                    help="states recorded per reactor")
    ap.add_argument("--sg-logged", type=int, default=1 << 20,
                    help="states logged from reacting-box runs")
    ap.add_argument("--sg-steps", type=int, default=4000,
                    help="Adam steps of the surrogate training")
    ap.add_argument("--ens-cases", type=int, default=48,
                    help="cases of the batched ensemble (0: skip)")
    ap.add_argument("--march-steps", type=int, default=100,
                    help="RK4 steps of the reacting-box marches")
    ap.add_argument("--out-dir", default=None,
                    help="directory for the ensemble tuning table")
    ap.add_argument("--mixture-kernels", type=int, default=0, metavar="REPS",
                    help="only evaluate the reacting-box right-hand sides (Euler and "
                         "Navier-Stokes, FP64 and mixed) REPS times each and exit "
                         "(a short run for kernel profiling)")
    return ap.parse_args()


def alg_bytes(K, np_, nfp, dim, nfld):
    """Algorithmic HBM bytes per launch of the two fused NS kernels: every
    nodal array read or written once, plus per-element geometry (rx,
    normals, face Jacobian ratios) and the int32 neighbour map once.
      mg_ns_flux: read q (F), write nodal net flux (F*dim)
      mg_ns_div:  read q (F) and nodal net flux (F*dim), write RHS (F)"""
    n = K * np_
    nf = dim + 1
    geo = K * (8 * (dim * dim + nf * dim + nf) + 4 * nf * nfp)
    flux = 8 * n * (nfld + nfld * dim) + geo
    div = 8 * n * (nfld + nfld * dim + nfld) + geo
    return {"mg_ns_flux": flux, "mg_ns_div": div}


def alg_flops(K, np_, nfp, dim, nfld):
    """Contraction flops per launch (element-local matrix applications)."""
    nfn = (dim + 1) * nfp
    flux = K * (nfld + 1) * (2 * dim * np_ * np_ + 2 * np_ * nfn)
    div = K * nfld * (2 * dim * np_ * np_ + 2 * np_ * nfn)
    return {"mg_ns_flux": flux, "mg_ns_div": div}


def cpu_baseline(n, order):
    """The FP64 oracle (NumPy) on a TGV mesh of n^3 x 6 tets, same order."""
    import numpy as np

    from oracle import initializers as oini, operators as oop, physics as ph
    from tests.cases import Problem

    cores = len(os.sched_getaffinity(0))
    pb = Problem(3, order, (n,) * 3, 0.0, 2 * np.pi, (True,) * 3)
    q0 = oini.taylor_green(pb.x)
    mu, kappa = _tgv_transport()
    ogm = ph.GasModel(ph.IdealSingleGas(),
                      ph.SimpleTransport(viscosity=mu, thermal_conductivity=kappa))
    reps = 0
    t0 = time.perf_counter()
    while True:
        st = ph.make_fluid_state(q0, ogm, 3)
        oop.ns_operator(pb.odisc, pb.box, st, ogm, {})
        reps += 1
        el = time.perf_counter() - t0
        if el > 10.0 or reps >= 20:
            break
    K = pb.mesh.nelements
    dofs = K * pb.dcoll.refel.nnodes * 5
    return {"value": dofs * reps / el, "unit": "DOF-RHS/s", "cores": cores,
            "kind": "port",
            "sample": f"FP64 NumPy oracle, TGV {n}^3x6={K} tets p={order}, "
                      f"{reps} RHS evaluations in {el:.1f} s "
                      f"(NumPy; BLAS threads <= {cores})"}


def measure_reacting(args, wl):
    """Plain-chemistry reacting box (uiuc_7sp, p = 2, n^3 x 6 tets, Euler +
    species sources, FP64): one RHS = Newton temperature from the previous
    stage's temperature + mixture Euler RHS + species sources added into it
    (mg_species_source_add)."""
    import torch

    from mirgecom_amd import _lib
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.workloads import reacting_box

    gm, dcoll = wl.gas_model, wl.dcoll
    n = round((dcoll.mesh.nelements / 6) ** (1 / 3))
    cv, tseed = wl.cv, wl.temperature
    K, npe = dcoll.mesh.nelements, dcoll.refel.nnodes
    nfld = cv.q.shape[0]
    import warnings

    def rhs():
        st = make_fluid_state(cv, gm, temperature_seed=tseed)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            r = euler_operator(dcoll, st, gm, {},
                               inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)
        return gm.eos.add_species_source_terms(cv, st.temperature, r)

    for _ in range(args.warmup):
        rhs()
    torch.cuda.synchronize()
    _lib.timing = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rhs()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    timing, _lib.timing = _lib.timing, None
    # the same state through the Navier-Stokes RHS with mixture-averaged
    # transport (fused FP64 path), timed on its own: not part of this RHS
    from mirgecom_amd.gas_model import GasModel
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import MixtureAveragedTransport
    gm_ns = GasModel(eos=gm.eos, transport=MixtureAveragedTransport())

    def ns_rhs():
        st = make_fluid_state(cv, gm_ns, temperature_seed=tseed)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            return ns_operator(dcoll, gm_ns, st, {},
                               inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)

    for _ in range(args.warmup):
        ns_rhs()
    torch.cuda.synchronize()
    _lib.timing = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ns_rhs()
    torch.cuda.synchronize()
    ms_ns = (time.perf_counter() - t0) / args.steps * 1e3
    timing.update({"ns_mixavg:" + k: v for k, v in _lib.timing.items()})
    _lib.timing = None
    # the species limiter of mixture drivers (mass fractions into [0, 1],
    # averages clipped), timed on its own: not part of this RHS
    from mirgecom_amd.limiter import bound_preserving_limiter
    y = cv.species_mass_fractions.contiguous()
    for _ in range(args.warmup):
        bound_preserving_limiter(dcoll, y, mmin=0.0, mmax=1.0, modify_average=True)
    _lib.timing = {}
    for _ in range(args.steps):
        bound_preserving_limiter(dcoll, y, mmin=0.0, mmax=1.0, modify_average=True)
    torch.cuda.synchronize()
    timing.update(_lib.timing)
    _lib.timing = None
    kern = {k: sum(a.elapsed_time(b) for a, b in v) / len(v)
            for k, v in timing.items()}
    nn = K * npe
    nfp = dcoll.refel.nface_nodes
    geo = K * (8 * (9 + 4 * 3 + 4) + 4 * 4 * nfp)
    nf = 3 * nfld
    ab = {"mg_mixture_temperature": 8 * nn * (nfld + 2),
          "mg_euler_rhs": 8 * nn * (2 * nfld + 1) + geo,
          "mg_species_source_add": 8 * nn * (1 + 7 + 1 + 7 + 7),
          "mg_bound_limiter": 8 * nn * 7 * 2,
          # mixture-averaged NS: gradients of the 12 fields and T written,
          # read back with the state for the node-parallel net flux
          "ns_mixavg:mg_mixture_temperature": 8 * nn * (nfld + 2),
          "ns_mixavg:mg_ns_flux": 8 * nn * (nfld + 1 + 2 * 3 * (nfld + 1) + nfld + 1 + nf)
                                   + geo,
          "ns_mixavg:mg_ns_div": 8 * nn * (nfld + 1 + nf + nfld) + geo}
    tab = {k: {"ms": round(kern[k], 4), "alg_bytes": ab.get(k),
               "GB_s": round(ab[k] / (kern[k] * 1e-3) / 1e9, 1) if k in ab else None}
           for k in kern}
    dofs = nn * nfld
    print(f"# reacting box {n}^3x6 p=2: {ms:.3f} ms/RHS, NS mixture-averaged "
          f"{ms_ns:.3f} ms/RHS, kernels {json.dumps(tab)}",
          file=sys.stderr, flush=True)
    return {"workload": f"3D reacting box uiuc_7sp, periodic, {n}^3x6={K} tets, "
                        "p=2, Euler + species sources (plain chemistry, FP64)",
            "dofs": dofs, "ms_per_rhs": ms, "value": dofs / (ms * 1e-3),
            "unit": "DOF-RHS/s", "kernels": tab,
            "ns_mixture_averaged": {
                "workload": "same state, Navier-Stokes RHS with "
                            "MixtureAveragedTransport (fused FP64)",
                "ms_per_rhs": ms_ns, "value": dofs / (ms_ns * 1e-3)}}


def measure_copy(dev):
    """Measured HBM bandwidth of a plain streaming device copy (1 GiB read +
    1 GiB written per copy), next to the 8 TB/s datasheet figure."""
    import torch
    src = torch.ones(128 * 1024 * 1024, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    reps = 10
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    gbs = 2 * src.numel() * 8 / (ms * 1e-3) / 1e9
    del src, dst
    return {"kernel": "torch device copy, 1 GiB", "ms": ms, "GB_s": gbs,
            "datasheet_GB_s": HBM_PEAK_GBS, "frac_of_datasheet": gbs / HBM_PEAK_GBS}


def vortex_cpu_line(cores):
    """2-D Euler isentropic vortex, 2,048 triangles, p = 3: the FP64 oracle's
    RHS timed on the host (the CPU path), the GPU RHS timed, and the L-inf
    error of the GPU march to t = 1 against the exact advected vortex (the
    GPU RHS equals the oracle's to 1e-12, tests/test_gpu_workloads.py)."""
    import numpy as np
    import torch

    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.graphs import GraphedStepper
    from mirgecom_amd.integrators import rk4_step
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.workloads import vortex_2d
    from oracle import initializers as oini, operators as oop, physics as ph
    from oracle.disc import OracleDisc

    wl, vortex = vortex_2d()
    gm = wl.gas_model
    m = wl.dcoll.mesh
    od = OracleDisc(m.vertices, m.elements, wl.dcoll.refel.nodes, wl.dcoll.order,
                    (m.a, m.b, m.periodic))
    box = (m.a, m.b, m.periodic)
    ogm = ph.GasModel(ph.IdealSingleGas())
    kw = dict(beta=5, center=(5, 0), velocity=(1, 0))
    obnd = {"all": oop.PrescribedFluidBoundary(
        lambda xb, sm, t: ph.make_fluid_state(oini.vortex2d(xb, t, **kw), ogm, 2))}
    q = wl.cv.q.cpu().numpy()
    reps = 0
    t0 = time.perf_counter()
    while True:
        oop.euler_operator(od, box, ph.make_fluid_state(q, ogm, 2), ogm, obnd, time=0.0)
        reps += 1
        el = time.perf_counter() - t0
        if el > 3.0 or reps >= 50:
            break
    cpu_ms = el / reps * 1e3

    def rhs(t, cv):
        return euler_operator(wl.dcoll, make_fluid_state(cv, gm), gm, wl.boundaries,
                              time=t, inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)

    for _ in range(3):
        rhs(0.0, wl.cv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        rhs(0.0, wl.cv)
    torch.cuda.synchronize()
    gpu_ms = (time.perf_counter() - t0) / 20 * 1e3
    cv = wl.cv
    nsteps = int(round(1.0 / wl.dt))
    for k in range(nsteps):
        cv = rk4_step(cv, k * wl.dt, wl.dt, rhs)
    exact = vortex(wl.dcoll.nodes(), time=1.0).q
    linf = float((cv.q - exact).abs().max())
    dofs = wl.dcoll.mesh.nelements * wl.dcoll.refel.nnodes * 4
    del GraphedStepper, np
    return {"workload": "2D Euler isentropic vortex, [0,10]x[-5,5], 32x32x2 = 2048 "
                        "triangles, p=3, Rusanov, exact prescribed boundaries",
            "cpu_rhs_ms": cpu_ms, "cpu_dof_rhs_s": dofs / (cpu_ms * 1e-3),
            "cpu_cores": cores, "cpu_kind": "FP64 NumPy oracle",
            "gpu_rhs_ms": gpu_ms, "gpu_dof_rhs_s": dofs / (gpu_ms * 1e-3),
            "march": f"RK4 dt={wl.dt} x {nsteps} to t=1 (GPU FP64)",
            "linf_vs_exact": linf}


def measure_reacting_variants(args, wl, model=None):
    """The reacting box in its Euler form (config of measure_reacting):
    plain FP64, mixed precision (mg_euler_rhs_mixed) with the exact sources,
    and mixed precision with the surrogate's sources; per-RHS times and the
    per-field error of a march of args.march_steps RK4 steps at dt = 1e-8
    against the FP64 plain march."""
    import warnings

    import torch

    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.integrators import rk4_step
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov

    gm, dcoll, dt = wl.gas_model, wl.dcoll, wl.dt
    npe = dcoll.refel.nnodes

    def make_rhs(precision, source):
        tbox = [wl.temperature.clone()]

        def rhs(t, cv):
            st = make_fluid_state(cv, gm, temperature_seed=tbox[0])
            tbox[0] = st.temperature
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                r = euler_operator(dcoll, st, gm, {}, precision=precision,
                                   inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)
            if source == "exact":
                return gm.eos.add_species_source_terms(cv, st.temperature, r)
            return model.species_source(cv, st.temperature, dt, npe, rhs=r)
        return rhs

    variants = [("fp64_exact", "float64", "exact"), ("mixed_exact", "mixed", "exact")]
    if model is not None:
        variants.append(("mixed_surrogate", "mixed", "surrogate"))
    out = {}
    dofs = dcoll.mesh.nelements * npe * 12
    for name, prec, src in variants:
        f = make_rhs(prec, src)
        for _ in range(args.warmup):
            f(0.0, wl.cv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f(0.0, wl.cv)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        out[name] = {"ms_per_rhs": ms, "dof_rhs_s": dofs / (ms * 1e-3)}
    finals = {}
    for name, prec, src in variants:
        f = make_rhs(prec, src)
        cv = wl.cv
        for k in range(args.march_steps):
            cv = rk4_step(cv, k * dt, dt, f)
        torch.cuda.synchronize()
        finals[name] = cv.q
    ref = finals["fp64_exact"]
    eos = gm.eos
    cref = ConservedVars(3, ref)
    tref = eos.temperature(cref, temperature_seed=wl.temperature)
    cmax = float(eos.sound_speed(cref, tref).max())
    rhomax = float(ref[0].max())
    for name in finals:
        if name == "fp64_exact":
            continue
        q = finals[name]
        cq = ConservedVars(3, q)
        tq = eos.temperature(cq, temperature_seed=wl.temperature)
        out[name]["march_rel_err"] = _rel_err_mix(q, ref, 3)
        out[name]["march_max_abs_dT_K"] = float((tq - tref).abs().max())
        out[name]["march_max_abs_dY"] = float(
            (cq.species_mass_fractions - cref.species_mass_fractions).abs().max())
        out[name]["march_momentum_err_over_rho_c"] = float(
            (q[2:5] - ref[2:5]).abs().max()) / (rhomax * cmax)
        out[name]["speedup_vs_fp64"] = (out["fp64_exact"]["ms_per_rhs"]
                                        / out[name]["ms_per_rhs"])
    res = {"workload": f"reacting box {dcoll.mesh.nelements} tets p=2, Euler + uiuc_7sp "
                       "sources, Newton T seeded from the previous stage",
           "dofs": dofs, "march": f"{args.march_steps} RK4 steps at dt={dt}", **out}
    print(f"# reacting Euler variants: {json.dumps(res)}", file=sys.stderr, flush=True)
    return res


def _ns_reacting_model(eos):
    """Constant transport of a 1600 K ethylene-air mixture (mu ~ 5e-5 Pa s,
    kappa ~ 0.1 W/(m K), d ~ 2.3e-4 m^2/s)."""
    import numpy as np

    from mirgecom_amd.gas_model import GasModel
    from mirgecom_amd.transport import SimpleTransport
    return GasModel(eos=eos, transport=SimpleTransport(
        viscosity=5e-5, thermal_conductivity=0.1,
        species_diffusivity=np.full(7, 2.3e-4)))


def measure_reacting_ns(args, wl, model=None):
    """The reacting box in Navier-Stokes form: plain FP64 (Newton + fused
    NS RHS + exact sources), mixed precision with the exact sources, and
    mixed precision with the surrogate's sources; per-RHS times and the
    per-field error of a march of args.march_steps RK4 steps at dt = 1e-8
    against the FP64 plain march."""
    import warnings

    import torch

    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.integrators import rk4_step
    from mirgecom_amd.navierstokes import ns_operator

    gm = _ns_reacting_model(wl.gas_model.eos)
    dcoll = wl.dcoll
    npe = dcoll.refel.nnodes
    dt = wl.dt

    def make_rhs(precision, source):
        tbox = [wl.temperature.clone()]

        def rhs(t, cv):
            st = make_fluid_state(cv, gm, temperature_seed=tbox[0])
            tbox[0] = st.temperature
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                r = ns_operator(dcoll, gm, st, {}, precision=precision)
            if source == "exact":
                return gm.eos.add_species_source_terms(cv, st.temperature, r)
            return model.species_source(cv, st.temperature, dt, npe, rhs=r)
        return rhs

    variants = [("fp64_exact", "float64", "exact"), ("mixed_exact", "mixed", "exact")]
    if model is not None:
        variants.append(("mixed_surrogate", "mixed", "surrogate"))
    out = {}
    nn = dcoll.mesh.nelements * npe
    dofs = nn * 12
    for name, prec, src in variants:
        f = make_rhs(prec, src)
        for _ in range(args.warmup):
            f(0.0, wl.cv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f(0.0, wl.cv)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        out[name] = {"ms_per_rhs": ms, "dof_rhs_s": dofs / (ms * 1e-3)}
    finals = {}
    from mirgecom_amd.fluid import ConservedVars
    for name, prec, src in variants:
        f = make_rhs(prec, src)
        cv = wl.cv
        for k in range(args.march_steps):
            cv = rk4_step(cv, k * dt, dt, f)
        torch.cuda.synchronize()
        finals[name] = cv.q
    ref = finals["fp64_exact"]
    eos = gm.eos
    cref = ConservedVars(3, ref)
    tref = eos.temperature(cref, temperature_seed=wl.temperature)
    cmax = float(eos.sound_speed(cref, tref).max())
    rhomax = float(ref[0].max())
    for name in finals:
        if name == "fp64_exact":
            continue
        q = finals[name]
        cq = ConservedVars(3, q)
        tq = eos.temperature(cq, temperature_seed=wl.temperature)
        out[name]["march_rel_err"] = _rel_err_mix(q, ref, 3)
        out[name]["march_max_abs_dT_K"] = float((tq - tref).abs().max())
        out[name]["march_max_abs_dY"] = float(
            (cq.species_mass_fractions - cref.species_mass_fractions).abs().max())
        out[name]["march_momentum_err_over_rho_c"] = float(
            (q[2:5] - ref[2:5]).abs().max()) / (rhomax * cmax)
        out[name]["speedup_vs_fp64"] = (out["fp64_exact"]["ms_per_rhs"]
                                        / out[name]["ms_per_rhs"])
    res = {"workload": f"reacting box {dcoll.mesh.nelements} tets p=2, Navier-Stokes "
                       "(constant transport) + uiuc_7sp sources, Newton T seeded from "
                       "the previous stage",
           "dofs": dofs, "march": f"{args.march_steps} RK4 steps at dt={dt}",
           **out}
    print(f"# reacting NS: {json.dumps(res)}", file=sys.stderr, flush=True)
    return res


def measure_surrogate(args, wl):
    """Train the chemistry surrogate on the GPU (autoignition trajectories +
    states logged from the reacting box) and evaluate it: held-out error
    (trained vs the same network untrained), ignition delays against the
    exact reactor, source error and fallback fraction on the reacting box,
    and the cost of its fused inference against the exact rates."""
    import numpy as np
    import torch

    from mirgecom_amd import _lib
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.surrogate import (
        ChemistrySurrogate, autoignition_dataset, premixed_ethylene_air,
        reactor_trajectories, states_from_field, substep_targets, surrogate_reactor)

    dt = 1e-8
    t0 = time.perf_counter()
    st_a, dy_a, _ = autoignition_dataset(args.sg_reactors, args.sg_samples, seed=7, dt=dt)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    # states logged from reacting-box RHS evaluations along a short march
    gm = wl.gas_model
    cv, tseed = wl.cv, wl.temperature
    logged = []
    per = args.sg_logged // 2
    for j in range(2):
        st = make_fluid_state(cv, gm, temperature_seed=tseed)
        logged.append(states_from_field(cv, st.temperature, dt, count=per, seed=j))
        from mirgecom_amd.euler import euler_operator
        from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
        import warnings
        for _ in range(10):
            st = make_fluid_state(cv, gm, temperature_seed=tseed)
            tseed = st.temperature
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                r = euler_operator(wl.dcoll, st, gm, {},
                                   inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)
            r = gm.eos.add_species_source_terms(cv, st.temperature, r)
            cv = cv + 5 * dt * r
    st_l = torch.cat(logged)
    dy_l = substep_targets(st_l, dt, 4)
    st_all = torch.cat([st_a, st_l])
    dy_all = torch.cat([dy_a, dy_l])
    del st_a, dy_a
    n = st_all.shape[0]
    g = torch.Generator(device="cpu").manual_seed(0)
    perm = torch.randperm(n, generator=g).to(st_all.device)
    nte = max(65536, n // 50)
    te, tr = perm[:nte], perm[nte:]
    model = ChemistrySurrogate(seed=0)
    model.fit_normalization(st_all[tr], dy_all[tr])
    x, tgt = model.features(st_all[tr], dy_all[tr])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model.train(x, tgt, steps=args.sg_steps, batch=8192, lr=2e-3, log_every=500)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del x, tgt
    untrained = ChemistrySurrogate(seed=0)
    untrained.norm = model.norm
    st_te, dy_te = st_all[te], dy_all[te]
    nte = (nte // 256) * 256
    st_te, dy_te = st_te[:nte], dy_te[:nte]
    scale = torch.tensor(list(model.norm.out_scale), dtype=torch.float64,
                         device=st_te.device)
    pred = model.predict_dy(st_te)
    e_tr = float(((pred - dy_te) / scale).pow(2).mean().sqrt())
    e_un = float(((untrained.predict_dy(st_te) - dy_te) / scale).pow(2).mean().sqrt())
    # per species over the norm of the largest species change (N2 is inert)
    dyn = torch.linalg.norm(dy_te, dim=0)
    l2_species = (torch.linalg.norm(pred - dy_te, dim=0) / dyn.max()).cpu().tolist()
    # ignition delays: exact reactor vs the network alone
    rng = np.random.default_rng(123)
    nr = 32
    t0r = rng.uniform(1400.0, 1600.0, nr)
    phi = rng.uniform(0.6, 1.6, nr)
    pr = rng.uniform(0.7, 1.5, nr) * 101325.0
    y0 = premixed_ethylene_air(phi)
    ti_ex, _, _ = reactor_trajectories(t0r, pr, y0, dt=dt, nmax=4000)
    ti_ex = ti_ex.cpu().numpy()
    nsteps = int(min(4000, 1.3 * ti_ex.max() / dt + 10))
    ti_sg, t_end_sg, _ = surrogate_reactor(model, t0r, pr, y0, nsteps, dt=dt)
    ti_sg = ti_sg.cpu().numpy()
    ok = (ti_ex > 0) & (ti_sg > 0)
    ign_err = np.abs(ti_sg - ti_ex)[ok] / ti_ex[ok]
    # on the reacting box: fallback fraction, source error, kernel times
    cvb = wl.cv
    stb = make_fluid_state(cvb, gm, temperature_seed=wl.temperature)
    tb = stb.temperature
    npe = wl.dcoll.refel.nnodes
    counter = torch.zeros(1, dtype=torch.int32, device=tb.device)
    src_sg = model.species_source(cvb, tb, dt, npe, fallback_counter=counter).q
    src_ex = gm.eos.get_species_source_terms(cvb, tb).q
    torch.cuda.synchronize()
    nel = wl.dcoll.mesh.nelements
    frac = int(counter) / nel
    sp = slice(5, 12)
    # over the norm of the largest species source (species absent from the
    # fresh mixture have no exact source)
    src_err = (torch.linalg.norm((src_sg[sp] - src_ex[sp]).reshape(7, -1), dim=1)
               / torch.linalg.norm(src_ex[sp].reshape(7, -1), dim=1).max())
    rhs = torch.zeros_like(cvb.q)
    from mirgecom_amd.fluid import ConservedVars
    rcv = ConservedVars(3, rhs)
    _lib.timing = {}
    for _ in range(args.steps):
        model.species_source(cvb, tb, dt, npe, rhs=rcv)
        gm.eos.add_species_source_terms(cvb, tb, rcv)
    torch.cuda.synchronize()
    kern = {k: sum(a.elapsed_time(b) for a, b in v) / len(v)
            for k, v in _lib.timing.items()}
    _lib.timing = None
    nn = nel * npe
    ms_sg = kern.get("mg_chem_surrogate_source")
    ms_ex = kern.get("mg_species_source_add")
    flop_pad = 2 * (32 * 256 + 2 * 256 * 256 + 256 * 32)
    flop_alg = 2 * (10 * 256 + 2 * 256 * 256 + 256 * 7)
    res = {"workload": "MLP 10->256->256->256->7 (GELU), BF16 on MFMA, FP32 "
                       "accumulation; trained on the GPU by the project's forward/"
                       "backward/Adam kernels",
           "data": {"autoignition_states": args.sg_reactors * args.sg_samples,
                    "reactors": args.sg_reactors,
                    "t0_K": [1000, 1600], "phi": [0.5, 2.0], "p_atm": [0.5, 2.0],
                    "logged_reacting_box_states": int(st_l.shape[0]),
                    "held_out": int(nte), "generation_s": t_data},
           "training": {"steps": args.sg_steps, "batch": 8192, "seconds": t_train,
                        "loss_history": model.history},
           "held_out_rms_err_output_scales": {"trained": e_tr, "untrained": e_un,
                                              "gain": e_un / e_tr},
           "held_out_l2_dY_err_per_species_over_largest": l2_species,
           "ignition_delay": {"reactors": nr, "t0_K": [1400, 1600],
                              "compared": int(ok.sum()),
                              "rel_err_mean": float(ign_err.mean()) if ok.any() else None,
                              "rel_err_max": float(ign_err.max()) if ok.any() else None,
                              "end_T_mean_K": float(t_end_sg.mean())},
           "reacting_box": {"fallback_fraction": frac,
                            "source_l2_err_per_species_over_largest":
                                src_err.cpu().tolist(),
                            "surrogate_source_ms": ms_sg, "exact_source_ms": ms_ex,
                            "speedup_vs_exact_rates": (ms_ex / ms_sg) if ms_sg else None,
                            "infer_TFLOP_s_padded": (flop_pad * nn / (ms_sg * 1e-3) / 1e12
                                                     if ms_sg else None),
                            "infer_TFLOP_s_algorithmic": (flop_alg * nn / (ms_sg * 1e-3)
                                                          / 1e12 if ms_sg else None),
                            "bf16_dense_peak_TFLOP_s": 2500.0}}
    print(f"# surrogate: {json.dumps(res)}", file=sys.stderr, flush=True)
    return res, model


def measure_ensemble(args, model):
    """48 (args.ens_cases) reacting Navier-Stokes cases of uneven size and
    order in shared launches: the launch shape autotuned, RK4 steps of the
    shared launches (graph-replayed and eager, the faster counted) against
    the same steps with per-case launches, and each case's RHS error against
    its FP64 exact-chemistry RHS."""
    import warnings

    import torch

    from mirgecom_amd.ensemble import ReactingEnsemble, autotune, measure, random_specs
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.navierstokes import ns_operator

    t0 = time.perf_counter()
    specs = random_specs(args.ens_cases, seed=0)
    ens = ReactingEnsemble(specs, surrogate=model)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    path = None
    if args.out_dir:
        os.makedirs(args.out_dir, exist_ok=True)
        path = os.path.join(args.out_dir, "ensemble_tuning.json")
    table = autotune(ens, reps=3, write=path)
    perf = measure(ens, steps=3)
    # per-case error of the batched RHS (mixed, surrogate) against FP64 exact
    y0, t0_ = ens.y.clone(), ens.temperature.clone()
    ens.rhs("y")
    torch.cuda.synchronize()
    errs, errs_prec = [], []
    for i in range(len(specs)):
        gm = ens.gas_models[i]
        cv = ConservedVars(3, ens.case_view(y0, i).clone())
        seed = t0_[int(ens.node_off[i]):int(ens.node_off[i + 1])].view_as(ens.case_t(i))
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            st = make_fluid_state(cv, gm, temperature_seed=seed.clone())
            r64 = ns_operator(ens.dcolls[i], gm, st, {})
            r32 = ns_operator(ens.dcolls[i], gm, st, {}, precision="mixed")
        r64 = gm.eos.add_species_source_terms(cv, st.temperature, r64)
        r32 = gm.eos.add_species_source_terms(cv, st.temperature, r32)
        errs.append(_rel_err_mix(ens.case_view(ens.k, i), r64.q, 3))
        errs_prec.append(_rel_err_mix(r32.q, r64.q, 3))
    res = {"workload": f"{len(specs)} periodic 3D Navier-Stokes + uiuc_7sp cases, "
                       "n^3x6 tets with n in [8, 28], p in {1..4}, Re in [500, 5000], "
                       "mixed precision, FP64 Newton and RK4, surrogate sources "
                       "(gated)",
           "cases": [(s.n, s.order, round(s.reynolds, 1)) for s in specs],
           "nodes": ens.total_nodes, "dofs": ens.dofs, "setup_s": t_setup,
           "tuning": table["best"], "tuning_trials": len(table["trials"]),
           "rhs_ms_tuned": table["best_ms"], **perf,
           "per_case_rhs_err_vs_fp64_exact": {
               "metric": "max over fields; species rows over the largest species row",
               "mixed_surrogate": [max(e) for e in errs],
               "mixed_exact_rates": [max(e) for e in errs_prec]}}
    print(f"# ensemble: {json.dumps(res)}", file=sys.stderr, flush=True)
    del ens
    torch.cuda.empty_cache()
    return res


def measure_overintegration(args, dcoll, gm, cv):
    """Euler RHS of the same state on the nodal and on the quadrature
    (degree 2p) discretization, and the entropy-stable flux-differencing
    operator on that quadrature; the Navier-Stokes RHS on the quadrature
    (gradients + divergence) and its entropy-stable split (the viscous
    terms on the quadrature plus the flux-differencing Euler RHS); FP64."""
    import warnings

    import torch

    from mirgecom_amd import _lib
    from mirgecom_amd.dof_desc import DISCR_TAG_BASE, DISCR_TAG_QUAD
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.inviscid import entropy_stable_inviscid_facial_flux_rusanov
    from mirgecom_amd.navierstokes import ns_operator
    st = make_fluid_state(cv, _inviscid_model(gm))
    st_ns = make_fluid_state(cv, gm)
    out = {}
    for tag in (DISCR_TAG_BASE, DISCR_TAG_QUAD, "esdg", "ns_quad", "ns_esdg"):
        def go():
            if tag in ("ns_quad", "ns_esdg"):
                with warnings.catch_warnings():
                    warnings.simplefilter("ignore")
                    return ns_operator(dcoll, gm, st_ns, {},
                                       use_esdg=tag == "ns_esdg",
                                       quadrature_tag=DISCR_TAG_QUAD)
            if tag == "esdg":
                return euler_operator(
                    dcoll, st, st.gas_model, {}, use_esdg=True,
                    inviscid_numerical_flux_func=entropy_stable_inviscid_facial_flux_rusanov,
                    quadrature_tag=DISCR_TAG_QUAD)
            return euler_operator(dcoll, st, st.gas_model, {},
                                  inviscid_numerical_flux_func=inviscid_facial_flux_rusanov,
                                  quadrature_tag=tag)
        for _ in range(args.warmup):
            go()
        torch.cuda.synchronize()
        _lib.timing = {}
        for _ in range(args.steps):
            go()
        torch.cuda.synchronize()
        t = {k: sum(a.elapsed_time(b) for a, b in v) / len(v)
             for k, v in _lib.timing.items()}
        _lib.timing = None
        out[tag] = t
    qo = dcoll.quadrature
    res = {"workload": "Euler RHS of the same state (Rusanov), nodal vs "
                       f"quadrature of degree {qo.degree} ({qo.nq} volume, "
                       f"{qo.nfq} points per face)",
           "nodal_ms": out[DISCR_TAG_BASE].get("mg_euler_rhs"),
           "quad_ms": out[DISCR_TAG_QUAD].get("mg_euler_rhs_quad"),
           "esdg_ms": out["esdg"].get("mg_euler_rhs_esdg"),
           "esdg_workload": "entropy-stable flux differencing on the same "
                            "quadrature (entropy-stable Rusanov faces)",
           "ns_quad_ms": {k: out["ns_quad"].get(k) for k in
                          ("mg_ns_gradient_quad", "mg_ns_rhs_quad")},
           "ns_esdg_ms": {k: out["ns_esdg"].get(k) for k in
                          ("mg_ns_gradient_quad", "mg_ns_rhs_quad",
                           "mg_euler_rhs_esdg")},
           "ns_workload": "Navier-Stokes RHS of the same state on the "
                          "quadrature (Rusanov); the entropy-stable split: "
                          "viscous terms on the quadrature + flux differencing"}
    print(f"# overintegration: {json.dumps(res)}", file=sys.stderr, flush=True)
    return res


def _inviscid_model(gm):
    """The gas model without its transport (for the Euler operator)."""
    from mirgecom_amd.gas_model import GasModel
    return GasModel(eos=gm.eos)


def measured_traffic(entry, n, order):
    """HBM bytes per launch of ``entry`` on the 24^3x6 p=3 workload from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated),
    or None for other workloads."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "ns_rhs_tgv24_p3_traffic.json")
    if (n, order) != (24, 3) or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["kernels"].get(entry, {}).get("traffic_bytes")


def _tgv_transport():
    from mirgecom_amd.initializers import TaylorGreenVortex
    return TaylorGreenVortex(mach=0.1).transport_for(
        reynolds=1600.0, prandtl=0.71, gamma=1.4, gas_const=287.1)


def _rel_err_mix(a, b, dim):
    """Per-field error of a mixture state or RHS: mass, energy as
    max|a-b|/max|b|; momentum components over the largest momentum
    component; the species rows over the largest species row (a species
    absent everywhere would otherwise divide by zero)."""
    out = _rel_err_cv(a, b, dim)
    sp = slice(2 + dim, b.shape[0])
    scale = float(b[sp].abs().max())
    for f in range(2 + dim, b.shape[0]):
        out[f] = float((a[f] - b[f]).abs().max()) / scale if scale > 0 else 0.0
    return out


def _rel_err_cv(a, b, dim):
    """Per-field max|a-b|/max|b|; momentum components normalised by the
    largest momentum component."""
    import torch
    out = []
    nf = b.shape[0]
    mscale = b[2:2 + dim].abs().max()
    for f in range(nf):
        scale = mscale if 2 <= f < 2 + dim else b[f].abs().max()
        out.append(float((a[f] - b[f]).abs().max() / scale))
    del torch
    return out


def mixture_kernels(args):
    """The reacting box's right-hand sides (32^3 x 6 tets, p = 2, uiuc_7sp):
    Euler FP64 and mixed, Navier-Stokes with constant transport FP64 and
    mixed, each evaluated args.mixture_kernels times; then, on the quadrature
    of degree 4, the overintegrated Navier-Stokes RHS and the entropy-stable
    Euler RHS (Renac's flux), each args.mixture_kernels times, with the
    per-entry-point HIP-event averages printed."""
    import warnings

    import torch

    from mirgecom_amd import _lib
    from mirgecom_amd.dof_desc import DISCR_TAG_QUAD
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.workloads import reacting_box

    torch.cuda.set_device(0)
    wl = reacting_box(n=args.reacting_n or 32, order=2)
    gm_ns = _ns_reacting_model(wl.gas_model.eos)
    ste = make_fluid_state(wl.cv, wl.gas_model, temperature_seed=wl.temperature)
    stn = make_fluid_state(wl.cv, gm_ns, temperature_seed=wl.temperature)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for prec in ("float64", "mixed"):
            for _ in range(args.mixture_kernels):
                euler_operator(wl.dcoll, ste, wl.gas_model, {}, precision=prec,
                               inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)
            for _ in range(args.mixture_kernels):
                ns_operator(wl.dcoll, gm_ns, stn, {}, precision=prec)
        torch.cuda.synchronize()
        _lib.timing = {}
        for _ in range(args.mixture_kernels):
            ns_operator(wl.dcoll, gm_ns, stn, {}, quadrature_tag=DISCR_TAG_QUAD)
        for _ in range(args.mixture_kernels):
            euler_operator(wl.dcoll, ste, wl.gas_model, {}, use_esdg=True,
                           quadrature_tag=DISCR_TAG_QUAD)
        torch.cuda.synchronize()
    t = {k: round(sum(a.elapsed_time(b) for a, b in v) / len(v), 4)
         for k, v in _lib.timing.items()}
    _lib.timing = None
    print(f"# mixture kernels: {args.mixture_kernels} evaluations per operator and "
          f"precision; quadrature-degree-4 entry points (ms): {json.dumps(t)}",
          file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU path is a design note, not built)",
              file=sys.stderr)
        sys.exit(2)
    if args.mixture_kernels:
        mixture_kernels(args)
        return
    import math

    import torch

    from mirgecom_amd import _lib
    from mirgecom_amd.discretization import (
        create_discretization_collection, generate_box_mesh)
    from mirgecom_amd.eos import IdealSingleGas
    from mirgecom_amd.gas_model import GasModel, make_fluid_state
    from mirgecom_amd.initializers import TaylorGreenVortex
    from mirgecom_amd.integrators import rk4_step
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.navierstokes import ns_operator
    from mirgecom_amd.transport import SimpleTransport

    n, order = args.n, args.order
    t0 = time.perf_counter()
    mesh = generate_box_mesh(3, 0.0, 2 * math.pi, (n,) * 3, periodic=(True,) * 3)
    dcoll = create_discretization_collection(mesh, order)
    print(f"# setup: {mesh.nelements} tets p={order} in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    tgv = TaylorGreenVortex(mach=0.1)
    mu, kappa = _tgv_transport()
    gm = GasModel(eos=IdealSingleGas(gamma=1.4, gas_const=287.1),
                  transport=SimpleTransport(viscosity=mu,
                                            thermal_conductivity=kappa))
    cv = tgv(dcoll.nodes(), eos=gm.eos)
    K, npe = mesh.nelements, dcoll.refel.nnodes
    nfld = 5
    dofs = K * npe * nfld

    def rhs(state_cv, precision):
        return ns_operator(dcoll, gm, make_fluid_state(state_cv, gm), {},
                           inviscid_numerical_flux_func=inviscid_facial_flux_rusanov,
                           precision=precision)

    def measure(precision):
        # warm up with the per-launch events on as well: the first few
        # hundred HIP events cost host time to create
        _lib.timing = {}
        for _ in range(max(args.warmup, 3 * args.steps // 2)):
            rhs(cv, precision)
        torch.cuda.synchronize()
        _lib.timing = {}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rhs(cv, precision)
        t_issue = time.perf_counter() - t0
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        print(f"# {precision}: host issue {t_issue / args.steps * 1e3:.3f} ms/step, "
              f"wall {wall / args.steps * 1e3:.3f} ms/step", file=sys.stderr,
              flush=True)
        timing = _lib.timing
        _lib.timing = None
        kern = {k: sum(a.elapsed_time(b) for a, b in v) / len(v)
                for k, v in timing.items()}
        return wall / args.steps * 1e3, kern

    ms64_eager, k64 = measure("float64")
    ms32_eager, k32 = measure("mixed")

    def measure_graph(precision):
        """The same RHS captured into a HIP graph and replayed (no host
        dispatch); the state buffer stays resident in HBM."""
        from mirgecom_amd.graphs import GraphedRHS
        g = GraphedRHS(lambda t, s: rhs(s, precision), cv)
        for _ in range(args.warmup):
            g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        err = float((g.out - rhs(cv, precision).q).abs().max())
        print(f"# {precision}: HIP-graph replay {ms:.3f} ms/step "
              f"(max |graph - eager| = {err:.1e})", file=sys.stderr, flush=True)
        del g
        return ms

    ms64_graph = ms32_graph = float("inf")
    if not args.no_graph:
        ms64_graph = measure_graph("float64")
        ms32_graph = measure_graph("mixed")
    ms64 = min(ms64_eager, ms64_graph)
    ms32 = min(ms32_eager, ms32_graph)

    # accuracy of the mixed path against FP64 on this workload: the RHS, and
    # the state after RK4 steps at CFL 0.3 (FP64 state and accumulation)
    err_rhs = _rel_err_cv(rhs(cv, "mixed").q, rhs(cv, "float64").q, 3)
    h = 2 * math.pi / n
    c0 = math.sqrt(1.4 * tgv.p0(1.4) / tgv.rho0)
    dt = 0.3 * h / (order**2 * (tgv.v0 + c0))
    nsteps = args.state_steps
    a = b = cv
    for k in range(nsteps):
        a = rk4_step(a, k * dt, dt, lambda t, s: rhs(s, "float64"))
        b = rk4_step(b, k * dt, dt, lambda t, s: rhs(s, "mixed"))
    err_state = _rel_err_cv(b.q, a.q, 3)
    mixed_ok = max(err_state) <= 1e-6
    print(f"# mixed vs fp64: rhs rel err {err_rhs}, state after {nsteps} RK4 "
          f"steps (dt={dt:.3e}) {err_state}", file=sys.stderr, flush=True)

    abytes = alg_bytes(K, npe, dcoll.refel.nface_nodes, 3, nfld)
    aflops = alg_flops(K, npe, dcoll.refel.nface_nodes, 3, nfld)
    # mixed path: nodal fluxes are FP32
    abytes32 = {"mg_ns_flux_mixed": abytes["mg_ns_flux"] - 4 * K * npe * nfld * 3,
                "mg_ns_div_mixed": abytes["mg_ns_div"] - 4 * K * npe * nfld * 3}
    aflops32 = {"mg_ns_flux_mixed": aflops["mg_ns_flux"],
                "mg_ns_div_mixed": aflops["mg_ns_div"]}

    def table(kern, ab, af):
        return {k: {"ms": round(kern[k], 4), "alg_bytes": ab[k],
                    "GB_s": round(ab[k] / (kern[k] * 1e-3) / 1e9, 1),
                    "TFLOP_s_contractions": round(af[k] / (kern[k] * 1e-3) / 1e12, 2)}
                for k in kern}

    t64 = table(k64, abytes, aflops)
    t32 = table(k32, abytes32, aflops32)
    print("# kernels fp64: " + json.dumps(t64), file=sys.stderr, flush=True)
    print("# kernels mixed: " + json.dumps(t32), file=sys.stderr, flush=True)
    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("# copy calibration: 3 x copy of 268435456 bytes", file=sys.stderr,
              flush=True)
    copy = measure_copy(dev)
    print(f"# copy: {json.dumps(copy)}", file=sys.stderr, flush=True)
    reacting = reacting_ns = surrogate = ensemble = None
    model = None
    if args.reacting_n:
        from mirgecom_amd.workloads import reacting_box
        wl3 = reacting_box(n=args.reacting_n, order=2)
        reacting = measure_reacting(args, wl3)
        if not args.no_surrogate:
            surrogate, model = measure_surrogate(args, wl3)
        reacting["precision_variants"] = measure_reacting_variants(args, wl3, model)
        reacting_ns = measure_reacting_ns(args, wl3, model)
        del wl3
    if args.ens_cases:
        ensemble = measure_ensemble(args, model)
    overint = measure_overintegration(args, dcoll, gm, cv)
    cpu = vortex_line = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_n, order)
        vortex_line = vortex_cpu_line(cpu["cores"])
        print(f"# vortex 2d: {json.dumps(vortex_line)}", file=sys.stderr, flush=True)

    ms = ms32 if mixed_ok else ms64
    kern, ab = (k32, abytes32) if mixed_ok else (k64, abytes)
    dom = max(kern, key=kern.get)
    achieved = ab[dom] / (kern[dom] * 1e-3) / 1e9
    dom64 = max(k64, key=k64.get)
    ach64 = abytes[dom64] / (k64[dom64] * 1e-3) / 1e9
    rhs64_bytes = sum(abytes[k] for k in k64)
    rhs64_ms = sum(k64.values())
    roof64 = {"bound": "hbm", "kernel": dom64, "achieved": ach64,
              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach64 / HBM_PEAK_GBS,
              "frac_of_measured_copy": ach64 / copy["GB_s"],
              "traffic": measured_traffic(dom64, n, order),
              "alg_bytes": abytes[dom64],
              "whole_rhs": {"alg_bytes": rhs64_bytes, "kernel_ms": rhs64_ms,
                            "GB_s": rhs64_bytes / (rhs64_ms * 1e-3) / 1e9,
                            "frac": rhs64_bytes / (rhs64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    out = {
        "metric": "DOF-RHS/s (3D NS p=3) mixed vs FP64 at <=1e-6 rel err; "
                  "chem-surrogate speedup",
        "value": dofs / (ms * 1e-3),
        "unit": "DOF-RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed (f64 state, physics, RK)" if mixed_ok else "f64",
        "data": "synthetic (Taylor-Green vortex initial state)",
        "config": {"workload": f"3D NS Taylor-Green, periodic [0,2pi]^3, "
                               f"{n}^3x6={K} tets, p={order}, Re=1600, "
                               f"Ma=0.1, Pr=0.71, Rusanov + central viscous",
                   "nodes": K * npe, "fields": nfld, "dofs": dofs,
                   "timing": "best of eager calls and HIP-graph replay of "
                             "the same RHS (both listed per precision)",
                   "fp64": {"value": dofs / (ms64 * 1e-3), "ms_per_step": ms64,
                            "eager_ms": ms64_eager,
                            "graph_ms": None if args.no_graph else ms64_graph,
                            "kernels": t64},
                   "mixed": {"value": dofs / (ms32 * 1e-3), "ms_per_step": ms32,
                             "eager_ms": ms32_eager,
                             "graph_ms": None if args.no_graph else ms32_graph,
                             "kernels": t32,
                             "speedup_vs_fp64": ms64 / ms32,
                             "rel_err_rhs": err_rhs,
                             "rel_err_state": err_state,
                             "state_steps": nsteps, "dt": dt},
                   "fp64_roofline": roof64,
                   "hbm_copy": copy,
                   "reacting": reacting,
                   "reacting_ns": reacting_ns,
                   "surrogate": surrogate,
                   "ensemble": ensemble,
                   "vortex_2d": vortex_line,
                   "overintegration": overint},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "frac_of_measured_copy": achieved / copy["GB_s"],
                     "traffic": measured_traffic(dom, n, order),
                     "alg_bytes": ab[dom]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
