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


def measure_reacting(args, n):
    """Plain-chemistry reacting box (uiuc_7sp, p = 2, n^3 x 6 tets, Euler +
    species sources, FP64): one RHS = Newton temperature from the previous
    stage's temperature + mixture Euler RHS + species source + their sum."""
    import torch

    from mirgecom_amd import _lib
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.workloads import reacting_box

    wl = reacting_box(n=n, order=2)
    gm, dcoll = wl.gas_model, wl.dcoll
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
        return r + gm.eos.get_species_source_terms(cv, st.temperature)

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
    ab = {"mg_mixture_temperature": 8 * nn * (nfld + 2),
          "mg_euler_rhs": 8 * nn * (2 * nfld + 1) + geo,
          "mg_species_source": 8 * nn * (1 + 7 + 1 + 7),
          "mg_bound_limiter": 8 * nn * 7 * 2}
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


def measure_overintegration(args, dcoll, gm, cv):
    """Euler RHS of the same state on the nodal and on the quadrature
    (degree 2p) discretization, and the entropy-stable flux-differencing
    operator on that quadrature, FP64."""
    import torch

    from mirgecom_amd import _lib
    from mirgecom_amd.dof_desc import DISCR_TAG_BASE, DISCR_TAG_QUAD
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.inviscid import entropy_stable_inviscid_facial_flux_rusanov
    st = make_fluid_state(cv, _inviscid_model(gm))
    out = {}
    for tag in (DISCR_TAG_BASE, DISCR_TAG_QUAD, "esdg"):
        def go():
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
                            "quadrature (entropy-stable Rusanov faces)"}
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU path is a design note, not built)",
              file=sys.stderr)
        sys.exit(2)
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
    reacting = measure_reacting(args, args.reacting_n) if args.reacting_n else None
    overint = measure_overintegration(args, dcoll, gm, cv)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_n, order)

    ms = ms32 if mixed_ok else ms64
    kern, ab = (k32, abytes32) if mixed_ok else (k64, abytes)
    dom = max(kern, key=kern.get)
    achieved = ab[dom] / (kern[dom] * 1e-3) / 1e9
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
                   "reacting": reacting,
                   "overintegration": overint},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, n, order),
                     "alg_bytes": ab[dom]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
