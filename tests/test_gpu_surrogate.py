"""The chemistry surrogate on the GPU (mirgecom_amd/csrc/mg_chem.hip):

* the constant-pressure reactor and substep kernels against the oracle
  reactor (oracle/reactor.py) at <= 1e-10;
* the BF16 matrix-core network: forward outputs and the gradients of the
  backward kernels against the FP64 oracle network (oracle/mlp.py) with the
  kernels' BF16 rounding emulated (<= 3e-3 relative, FP32 accumulation
  order) and without it (<= 5e-2, the BF16 rounding itself); one Adam step
  against its formula;
* training on autoignition trajectories lowers the loss, the fused gated
  inference equals the training forward pass, the gate sends out-of-range
  and failed-check elements to the exact rates (bit-equal to
  mg_species_source), and the trained network's gain over the same network
  untrained and its ignition-delay error against the exact reactor.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _premixed(phi):
    from mirgecom_amd.surrogate import premixed_ethylene_air
    return premixed_ethylene_air(np.asarray(phi, dtype=float))


def test_reactor_against_oracle():
    from mirgecom_amd.surrogate import reactor_trajectories
    from oracle import reactor

    print("reactor kernel vs oracle, 4 reactors x 600 steps", flush=True)
    t0 = np.array([1600.0, 1500.0, 1550.0, 1450.0])
    phi = np.array([0.7, 1.0, 1.3, 1.8])
    p = np.array([0.8, 1.0, 1.5, 0.6]) * 101325.0
    y0 = _premixed(phi)
    nmax, nsamp = 600, 6
    _, st, dy = reactor_trajectories(t0, p, y0, nmax=nmax, t_rise=1e9, nsamp=nsamp)
    st = st.cpu().numpy().reshape(nsamp, 4, 10)
    dy = dy.cpu().numpy().reshape(nsamp, 4, 7)
    t, y = t0.copy(), y0.copy()
    from oracle import thermochem as tc
    h0 = tc.mixture_enthalpy_mass(t, y)
    for s in range(nmax):
        if s % (nmax // nsamp) == 0:
            j = s // (nmax // nsamp)
            assert np.abs(st[j, :, 0] - t).max() <= 1e-10 * t.max()
            assert np.abs(st[j, :, 2:9] - y.T).max() <= 1e-10
        tn, yn = reactor.rk4_step(p, h0, 1e-8, t, y)
        if s % (nmax // nsamp) == 0:
            ref = (yn - y).T
            assert np.abs(dy[j] - ref).max() <= 1e-10 * np.abs(ref).max() + 1e-18
        t, y = tn, yn
    # ignition delays: same step count
    ti, _, _ = reactor_trajectories(t0[:2], p[:2], y0[:, :2], nmax=3000)
    _, _, ti_ref = reactor.march(t0[:2], p[:2], y0[:, :2], 1e-8, 3000)
    assert np.all(ti_ref > 0)
    assert np.abs(ti.cpu().numpy() - ti_ref).max() <= 1.5e-8


def test_substep_against_oracle():
    from mirgecom_amd.surrogate import substep_targets
    from oracle import reactor
    rng = np.random.default_rng(3)
    n = 256
    y = _premixed(rng.uniform(0.5, 2.0, n))
    y[[2, 3, 4, 5]] += rng.uniform(0, 0.02, (4, n))
    y /= y.sum(axis=0)
    st = np.concatenate([rng.uniform(1100, 2400, (n, 1)),
                         rng.uniform(0.5, 2, (n, 1)) * 101325.0, y.T,
                         np.full((n, 1), 1e-8)], 1)
    got = substep_targets(torch.as_tensor(st, device="cuda"), 1e-8, 4).cpu().numpy()
    ref = reactor.substep(st, 1e-8, 4)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def _random_batch(model, b, seed):
    from oracle import mlp
    rng = np.random.default_rng(seed)
    x = np.zeros((b, 32))
    x[:, :10] = rng.normal(size=(b, 10))
    x = mlp.to_bf16(x)
    tgt = np.zeros((b, 32))
    tgt[:, :7] = rng.normal(size=(b, 7))
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda").to(torch.bfloat16)
    tt = torch.as_tensor(tgt, dtype=torch.float32, device="cuda")
    return x, tgt, xt, tt


def _perturb_biases(model, seed):
    """Nonzero biases so that the bias paths are exercised."""
    from mirgecom_amd.surrogate import PARAM_OFFSETS, _LAYERS
    g = torch.Generator(device="cpu").manual_seed(seed)
    p = model.params.cpu()
    for (woff, boff), (kin, nout) in zip(PARAM_OFFSETS, _LAYERS):
        nreal = 7 if nout == 32 else nout
        p[boff:boff + nreal] = 0.1 * torch.randn(nreal, generator=g)
    model.params.copy_(p.to(model.device))
    model.sync_bf16()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_forward_against_oracle():
    from mirgecom_amd.surrogate import ChemistrySurrogate
    from oracle import mlp
    model = ChemistrySurrogate(seed=1)
    _perturb_biases(model, 2)
    x, _, xt, _ = _random_batch(model, 512, 4)
    out = model.forward(xt).cpu().numpy().astype(np.float64)
    layers = model.layer_weights()
    ref_bf, _ = mlp.forward(layers, x, bf16=True)
    ref64, _ = mlp.forward(layers, x, bf16=False)
    e_bf, e_64 = _rel(out[:, :7], ref_bf[:, :7]), _rel(out[:, :7], ref64[:, :7])
    print("forward rel err: bf16-emulated", e_bf, "fp64", e_64, flush=True)
    assert e_bf < 3e-3 and e_64 < 5e-2
    assert np.all(out[:, 7:] == 0.0)


def test_gradient_against_oracle():
    """dL/dparams of the forward + backward kernels against the oracle's
    reverse mode, layer by layer."""
    from mirgecom_amd.surrogate import PARAM_OFFSETS, ChemistrySurrogate
    from oracle import mlp
    model = ChemistrySurrogate(seed=5)
    _perturb_biases(model, 6)
    x, tgt, xt, tt = _random_batch(model, 1024, 7)
    loss, grad = model.gradient(xt, tt)
    grad = grad.cpu().numpy().astype(np.float64)
    layers = model.layer_weights()
    for bf16, tol in ((True, 3e-3), (False, 5e-2)):
        lref, gref = mlp.loss_and_grad(layers, x, tgt, bf16=bf16)
        assert abs(loss - lref) <= tol * lref
        for li, (woff, boff) in enumerate(PARAM_OFFSETS):
            kin, nout = gref[li][0].shape
            gw = grad[woff:boff].reshape(kin, nout)
            gb = grad[boff:boff + nout]
            ew, eb = _rel(gw, gref[li][0]), _rel(gb, gref[li][1])
            print(f"grad layer {li} bf16={bf16}: W {ew:.2e} b {eb:.2e}", flush=True)
            assert ew < tol and eb < tol, (li, bf16)


def test_adam_step():
    from mirgecom_amd.surrogate import ChemistrySurrogate
    model = ChemistrySurrogate(seed=9)
    _, _, xt, tt = _random_batch(model, 256, 10)
    p0 = model.params.clone()
    _, grad = model.gradient(xt, tt)
    work = model._work(256)
    from mirgecom_amd import _lib
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    for step in (1, 2):
        _lib.call("mg_mlp_adam", _lib.ptr(model.params), _lib.ptr(model.m),
                  _lib.ptr(model.v), _lib.ptr(work), 256, step, lr, b1, b2, eps,
                  _lib.ptr(model.wbf16), None, _lib.stream_handle())
    g = grad.double()
    m = (1 - b1) * g
    v = (1 - b2) * g * g
    p = p0.double() - lr * (m / (1 - b1)) / (torch.sqrt(v / (1 - b2)) + eps)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1**2)) / (torch.sqrt(v / (1 - b2**2)) + eps)
    assert float((model.params.double() - p).abs().max()) < 1e-6
    # the BF16 copies follow the master weights
    from mirgecom_amd.surrogate import PARAM_OFFSETS
    woff, boff = PARAM_OFFSETS[1]
    w = model.params[woff:boff].reshape(256, 256)
    wt = model.wbf16[8192:8192 + 65536].reshape(256, 256)       # Wt_2 [N][K]
    assert torch.equal(wt.float(), w.T.to(torch.bfloat16).float())


@pytest.fixture(scope="module")
def trained():
    """A network trained on a small autoignition data set (T0 in
    [1300, 1600] K so that trajectories are short) plus the held-out set."""
    from mirgecom_amd.surrogate import ChemistrySurrogate, autoignition_dataset
    print("surrogate: data generation", flush=True)
    st, dy, info = autoignition_dataset(384, 64, seed=11, nmax=12_000,
                                        t0_range=(1300.0, 1600.0))
    rng = np.random.default_rng(0)
    perm = torch.as_tensor(rng.permutation(st.shape[0]), device="cuda")
    st, dy = st[perm], dy[perm]
    ntr = int(0.9 * st.shape[0])
    model = ChemistrySurrogate(seed=0)
    model.fit_normalization(st[:ntr], dy[:ntr])
    x, tgt = model.features(st[:ntr], dy[:ntr])
    untrained = ChemistrySurrogate(seed=0)
    untrained.norm = model.norm
    print("surrogate: training", flush=True)
    model.train(x, tgt, steps=800, batch=4096, lr=2e-3)
    print("surrogate: loss history", model.history, flush=True)
    return model, untrained, st[ntr:], dy[ntr:]


def test_training_lowers_loss(trained):
    model = trained[0]
    first, last = model.history[0][1], model.history[-1][1]
    assert np.isfinite(last) and last < 0.05 * first


def test_fused_inference_equals_forward(trained):
    """mg_chem_surrogate_source (gate off) gives rho dY_net / dt with dY_net
    from the training forward pass on the same states."""
    from mirgecom_amd import mechanisms as mech
    from mirgecom_amd.fluid import ConservedVars
    model, _, st, _ = trained
    n = 640
    st = st[:n]
    saved = [(model.norm.lo[i], model.norm.hi[i]) for i in range(10)]
    period = model.norm.sample_period
    for i in range(10):
        model.norm.lo[i], model.norm.hi[i] = -1e300, 1e300
    model.norm.sample_period = 0
    try:
        t, p, y = st[:, 0], st[:, 1], st[:, 2:9].T
        w = torch.as_tensor(mech.molecular_weights(), dtype=t.dtype, device=t.device)
        rho = p / (mech.GAS_CONSTANT * (y / w[:, None]).sum(0) * t)
        q = torch.cat([rho[None], torch.zeros((4, n), dtype=t.dtype, device=t.device),
                       rho[None] * y], 0)
        cv = ConservedVars(3, q.reshape(12, n // 10, 10).contiguous())
        src = model.species_source(cv, t.reshape(n // 10, 10), 1e-8, 10).q
        ref = model.predict_dy(st) * rho[:, None] / 1e-8
        scale = torch.tensor(list(model.norm.out_scale), dtype=torch.float64,
                             device="cuda") * rho[:, None] / 1e-8
        err = ((src[5:].reshape(7, n).T - ref) / scale).abs().max()
        print("fused vs forward, in output scales:", float(err), flush=True)
        assert float(err) < 1e-3
    finally:
        for i in range(10):
            model.norm.lo[i], model.norm.hi[i] = saved[i]
        model.norm.sample_period = period


def _field_state(st, n):
    from mirgecom_amd import mechanisms as mech
    from mirgecom_amd.fluid import ConservedVars
    t, p, y = st[:n, 0], st[:n, 1], st[:n, 2:9].T
    w = torch.as_tensor(mech.molecular_weights(), dtype=t.dtype, device=t.device)
    rho = p / (mech.GAS_CONSTANT * (y / w[:, None]).sum(0) * t)
    q = torch.cat([rho[None], torch.zeros((4, n), dtype=t.dtype, device=t.device),
                   rho[None] * y], 0)
    return ConservedVars(3, q.reshape(12, n // 10, 10).contiguous()), \
        t.reshape(n // 10, 10).contiguous()


def test_gate_fallback_is_exact(trained):
    from mirgecom_amd.eos import PyrometheusMixture
    model, _, st, _ = trained
    n = 1280
    cv, t = _field_state(st, n)
    exact = PyrometheusMixture().get_species_source_terms(cv, t).q
    counter = torch.zeros(1, dtype=torch.int32, device="cuda")
    # every node out of range (T far below the training set)
    cold = t * 0.3
    exact_cold = PyrometheusMixture().get_species_source_terms(cv, cold).q
    got = model.species_source(cv, cold, 1e-8, 10, fallback_counter=counter).q
    assert torch.equal(got, exact_cold)
    assert int(counter) == n // 10
    # zero tolerance on a check of every node: every element falls back
    tol, period, off = model.norm.check_tol, model.norm.sample_period, \
        model.norm.sample_offset
    model.norm.check_tol, model.norm.sample_period, model.norm.sample_offset = 0.0, 1, 0
    try:
        counter.zero_()
        got = model.species_source(cv, t, 1e-8, 10, fallback_counter=counter).q
        assert torch.equal(got, exact) and int(counter) == n // 10
    finally:
        model.norm.check_tol = tol
        model.norm.sample_period, model.norm.sample_offset = period, off
    # in-distribution with the trained gate: most elements keep the network
    counter.zero_()
    model.species_source(cv, t, 1e-8, 10, fallback_counter=counter)
    frac = int(counter) / (n // 10)
    print("in-distribution fallback fraction", frac, flush=True)
    assert frac < 0.5


def test_gain_of_training(trained):
    """What the trained network gains over the same network untrained, on
    held-out trajectory states (error of dY in output scales), and its
    ignition delay against the exact reactor."""
    from mirgecom_amd.surrogate import reactor_trajectories, surrogate_reactor
    model, untrained, st, dy = trained
    scale = torch.tensor(list(model.norm.out_scale), dtype=torch.float64,
                         device="cuda")
    n = (st.shape[0] // 256) * 256
    e_tr = ((model.predict_dy(st[:n]) - dy[:n]) / scale).pow(2).mean().sqrt()
    e_un = ((untrained.predict_dy(st[:n]) - dy[:n]) / scale).pow(2).mean().sqrt()
    print(f"held-out RMS dY error (output scales): trained {float(e_tr):.3e}, "
          f"untrained {float(e_un):.3e}", flush=True)
    assert float(e_tr) < 0.1 * float(e_un)
    t0 = np.array([1600.0, 1550.0, 1500.0, 1600.0])
    phi = np.array([1.0, 0.8, 1.2, 1.5])
    p = np.array([1.0, 1.2, 0.9, 1.0]) * 101325.0
    y0 = _premixed(phi)
    ti_ex, _, _ = reactor_trajectories(t0, p, y0, nmax=4000)
    ti_sg, _, _ = surrogate_reactor(model, t0, p, y0, 2500)
    ti_ex, ti_sg = ti_ex.cpu().numpy(), ti_sg.cpu().numpy()
    print("ignition delay exact", ti_ex, "surrogate", ti_sg, flush=True)
    assert np.all(ti_ex > 0) and np.all(ti_sg > 0)
    assert np.abs(ti_sg - ti_ex).max() <= 0.25 * ti_ex.max()
