"""Learned chemistry surrogate of the uiuc_7sp mixture on the GPU.

A 10 -> 256 -> 256 -> 256 -> 7 GELU network maps (log T, log p, Y_1..Y_7,
dt) to the change of the mass fractions over a chemistry substep dt; the
species source is then ``S_k = rho dY_k / dt``.  It stands in for the exact
rates of ``get_species_source_terms`` (mirgecom/eos.py:858-875, rates of
mirgecom/mechanisms/uiuc_7sp.yaml:146-154) behind a gate that sends every
element with an out-of-range node, or a sampled node where the network misses
the exact rates, back to the exact rates.

Everything runs in libmgdg (include/mgdg.h, ``mirgecom_amd/csrc/mg_chem.hip``):

* training data: constant-pressure autoignition trajectories of the same
  mechanism (``mg_reactor_trajectories``; the reactor of
  test/test_chemistry.py:51-110) with T0 in [1000, 1600] K, equivalence ratio
  in [0.5, 2] and p in [0.5, 2] atm, plus states taken from reacting-box runs
  whose targets come from ``mg_chem_substep``;
* training: ``mg_mlp_forward`` / ``mg_mlp_backward`` / ``mg_mlp_adam`` (BF16
  on the matrix cores, FP32 accumulation and master weights);
* inference: ``mg_chem_surrogate_source`` (fused network, gate and fallback,
  optionally added into a right-hand side).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from mirgecom_amd import _lib
from mirgecom_amd import mechanisms as mech

NPARAM = 148256
NBF16 = 294912
IN_PAD, HIDDEN, OUT_PAD = 32, 256, 32
N_IN, N_OUT = 10, 7
# offsets of the FP32 master parameters (W_l [K_in][N_out], then b_l)
_LAYERS = [(IN_PAD, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, OUT_PAD)]


def _param_offsets():
    off, out = 0, []
    for kin, nout in _LAYERS:
        out.append((off, off + kin * nout))
        off += kin * nout + nout
    assert off == NPARAM
    return out


PARAM_OFFSETS = _param_offsets()


def premixed_ethylene_air(phi):
    """Mass fractions of ethylene / air (O2 + 3.76 N2 by mole) at
    equivalence ratio ``phi`` (array), species order of the mechanism."""
    phi = np.asarray(phi, dtype=np.float64)
    w = mech.molecular_weights()
    sp = {s: i for i, s in enumerate(mech.SPECIES)}
    x = np.zeros((7,) + phi.shape)
    x[sp["C2H4"]] = phi
    x[sp["O2"]] = 3.0
    x[sp["N2"]] = 3.0 * 3.76
    y = x * w.reshape((7,) + (1,) * phi.ndim)
    return y / y.sum(axis=0)


def _dev64(a, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)


def reactor_trajectories(t0, p, y0, dt=1e-8, nmax=120_000, t_rise=400.0, nsamp=0,
                         device="cuda"):
    """Run ``len(t0)`` constant-pressure reactors on the device; returns
    ``(t_ign [n], states [nsamp*n, 10], dy [nsamp*n, 7])`` (the last two None
    without samples)."""
    _lib.ensure_mechanism()
    t0 = _dev64(t0, device)
    p = _dev64(p, device)
    y0 = _dev64(y0, device)
    n = t0.numel()
    t_ign = torch.empty(n, dtype=torch.float64, device=device)
    states = dy = None
    if nsamp > 0:
        states = torch.empty((nsamp * n, 10), dtype=torch.float64, device=device)
        dy = torch.empty((nsamp * n, 7), dtype=torch.float64, device=device)
    _lib.call("mg_reactor_trajectories", n, _lib.ptr(t0), _lib.ptr(p), _lib.ptr(y0),
              float(dt), int(nmax), float(t_rise), int(nsamp), _lib.ptr(t_ign),
              _lib.ptr(states), _lib.ptr(dy), _lib.stream_handle())
    return t_ign, states, dy


def autoignition_dataset(n_reactors, nsamp, seed=0, dt=1e-8, nmax=120_000,
                         t0_range=(1000.0, 1600.0), device="cuda"):
    """States along seeded autoignition trajectories: T0 ~ U[t0_range] K
    (default 1000-1600 K), phi ~ U[0.5, 2], p ~ U[0.5, 2] atm;
    ``n_reactors * nsamp`` states with the change of Y over the next
    ``dt``."""
    rng = np.random.default_rng(seed)
    t0 = rng.uniform(t0_range[0], t0_range[1], n_reactors)
    phi = rng.uniform(0.5, 2.0, n_reactors)
    p = rng.uniform(0.5, 2.0, n_reactors) * 101325.0
    y0 = premixed_ethylene_air(phi)
    t_ign, states, dy = reactor_trajectories(t0, p, y0, dt=dt, nmax=nmax, nsamp=nsamp,
                                             device=device)
    return states, dy, {"t0": t0, "phi": phi, "p": p, "t_ign": t_ign}


def states_from_field(cv, temperature, dt, count=None, seed=0):
    """Surrogate input states ``[n, 10]`` = (T, p, Y, dt) of a mixture
    ``ConservedVars`` at ``temperature`` (all nodes, or ``count`` of them
    drawn without replacement)."""
    q = cv.q.reshape(cv.q.shape[0], -1)
    t = temperature.reshape(-1)
    rho = q[0]
    y = q[2 + cv.dim:] / rho
    w = torch.as_tensor(mech.molecular_weights(), dtype=q.dtype, device=q.device)
    rgas = mech.GAS_CONSTANT * (y / w[:, None]).sum(0)
    p = rho * rgas * t
    st = torch.cat([t[None], p[None], y, torch.full_like(t, dt)[None]], 0).T
    if count is not None and count < st.shape[0]:
        g = torch.Generator(device="cpu").manual_seed(seed)
        idx = torch.randperm(st.shape[0], generator=g)[:count].to(q.device)
        st = st[idx]
    return st.contiguous()


def substep_targets(states, dt=1e-8, nsub=4):
    """Change of Y over ``dt`` at constant pressure from each state."""
    _lib.ensure_mechanism()
    states = states.contiguous()
    dy = torch.empty((states.shape[0], 7), dtype=torch.float64, device=states.device)
    _lib.call("mg_chem_substep", states.shape[0], _lib.ptr(states), float(dt),
              int(nsub), _lib.ptr(dy), _lib.stream_handle())
    return dy


def _enthalpy_cp(t, y):
    """Mixture enthalpy and cp (mass basis) of ``y (7, n)`` at ``t (n,)``."""
    from mirgecom_amd.eos import _nasa7
    lo, hi = mech.nasa7_tables()
    cp_r, h_rt, _ = _nasa7(t, lo, hi)
    w = torch.as_tensor(mech.molecular_weights(), dtype=t.dtype,
                        device=t.device)[:, None]
    h = mech.GAS_CONSTANT * t * torch.sum(h_rt * y / w, dim=0)
    cp = mech.GAS_CONSTANT * torch.sum(cp_r * y / w, dim=0)
    return h, cp


def surrogate_reactor(model, t0, p, y0, nsteps, dt=1e-8, t_rise=400.0,
                      device="cuda"):
    """Constant-pressure reactors advanced by the network alone: each step
    ``Y += dY_net(T, p, Y, dt)``, T from the conserved enthalpy (4 Newton
    steps).  Returns ``(t_ign, T, Y)`` (t_ign = -1 without ignition)."""
    t = _dev64(t0, device).clone()
    p = _dev64(p, device)
    y = _dev64(y0, device).clone()
    t_init = t.clone()
    h0, _ = _enthalpy_cp(t, y)
    t_ign = torch.full_like(t, -1.0)
    for s in range(nsteps):
        st = torch.cat([t[None], p[None], y, torch.full_like(t, dt)[None]], 0).T
        y = y + model.predict_dy(st.contiguous()).T
        for _ in range(4):
            h, cp = _enthalpy_cp(t, y)
            t = t + (h0 - h) / cp
        new = (t_ign < 0) & (t > t_init + t_rise)
        t_ign = torch.where(new, torch.full_like(t, (s + 1) * dt), t_ign)
    return t_ign, t, y


class ChemistrySurrogate:
    """Parameters, normalisation and gate of the network, with training and
    inference through libmgdg."""

    def __init__(self, device="cuda", seed=0):
        self.device = torch.device(device)
        self.params = torch.zeros(NPARAM, dtype=torch.float32, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.wbf16 = torch.zeros(NBF16, dtype=torch.bfloat16, device=self.device)
        self.norm = _lib.MgChemNorm()
        self.step = 0
        self.history = []
        self._init_params(seed)

    # {{{ parameters

    def _init_params(self, seed):
        g = torch.Generator(device="cpu").manual_seed(seed)
        p = torch.zeros(NPARAM, dtype=torch.float32)
        real = [(N_IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_OUT)]
        for (woff, boff), (kin, nout), (kr, nr) in zip(PARAM_OFFSETS, _LAYERS, real):
            w = torch.zeros(kin, nout)
            w[:kr, :nr] = torch.randn(kr, nr, generator=g) * math.sqrt(2.0 / kr)
            p[woff:woff + kin * nout] = w.reshape(-1)
        self.params.copy_(p.to(self.device))
        self.sync_bf16()

    def sync_bf16(self):
        _lib.call("mg_mlp_sync_bf16", _lib.ptr(self.params), _lib.ptr(self.wbf16),
                  _lib.stream_handle())

    def layer_weights(self, params=None):
        """Host FP64 copies ``[(W [K_in][N_out], b [N_out]), ...]``."""
        p = (self.params if params is None else params).double().cpu().numpy()
        out = []
        for (woff, boff), (kin, nout) in zip(PARAM_OFFSETS, _LAYERS):
            out.append((p[woff:boff].reshape(kin, nout), p[boff:boff + nout]))
        return out

    # }}}

    # {{{ normalisation and gate

    def fit_normalization(self, states, dy, bound_margin=0.02, sample_period=100,
                          sample_offset=37, check_tol=0.5):
        raw = self._raw(states)
        mean = raw.mean(0)
        std = raw.std(0)
        scale = torch.where(std > 1e-12 * mean.abs().clamp_min(1e-300), std,
                            mean.abs().clamp_min(1e-300))
        lo = raw.min(0).values
        hi = raw.max(0).values
        span = (hi - lo).clamp_min(1e-12)
        omean = dy.mean(0)
        ostd = dy.std(0)
        oscale = torch.maximum(ostd, 1e-3 * ostd.max())
        nm = self.norm
        for i in range(N_IN):
            nm.in_mean[i] = float(mean[i])
            nm.in_scale[i] = float(scale[i])
            nm.lo[i] = float(lo[i] - bound_margin * span[i])
            nm.hi[i] = float(hi[i] + bound_margin * span[i])
        # dt enters as given; the bounds admit the trained substep only
        nm.lo[9], nm.hi[9] = float(lo[9]) * (1 - 1e-9), float(hi[9]) * (1 + 1e-9)
        for k in range(N_OUT):
            nm.out_mean[k] = float(omean[k])
            nm.out_scale[k] = float(oscale[k])
        nm.check_tol = float(check_tol)
        nm.sample_period = int(sample_period)
        nm.sample_offset = int(sample_offset)

    @staticmethod
    def _raw(states):
        return torch.cat([torch.log(states[:, 0:1]), torch.log(states[:, 1:2]),
                          states[:, 2:9], states[:, 9:10]], 1)

    def features(self, states, dy=None):
        """BF16 inputs ``[n, 32]`` (and FP32 normalised targets ``[n, 32]``)."""
        n = states.shape[0]
        states = states.contiguous()
        x = torch.empty((n, IN_PAD), dtype=torch.bfloat16, device=self.device)
        tgt = None
        if dy is not None:
            dy = dy.contiguous()
            tgt = torch.empty((n, OUT_PAD), dtype=torch.float32, device=self.device)
        _lib.call("mg_chem_features", n, _lib.ptr(states), _lib.ptr(dy), self.norm,
                  _lib.ptr(x), _lib.ptr(tgt), _lib.stream_handle())
        return x, tgt

    # }}}

    # {{{ training

    def forward(self, x, work=None):
        """Normalised outputs ``[B, 32]`` of a batch (B a multiple of 256)."""
        b = x.shape[0]
        work = self._work(b) if work is None else work
        out = torch.empty((b, OUT_PAD), dtype=torch.float32, device=self.device)
        _lib.call("mg_mlp_forward", b, _lib.ptr(x), _lib.ptr(self.params),
                  _lib.ptr(self.wbf16), _lib.ptr(work), _lib.ptr(out),
                  _lib.stream_handle())
        return out

    def gradient(self, x, tgt):
        """``(loss, dL/dparams)`` of one batch (no update)."""
        b = x.shape[0]
        work = self._work(b)
        _lib.call("mg_mlp_forward", b, _lib.ptr(x), _lib.ptr(self.params),
                  _lib.ptr(self.wbf16), _lib.ptr(work), None, _lib.stream_handle())
        lp = torch.empty((b * OUT_PAD + 255) // 256, dtype=torch.float32,
                         device=self.device)
        _lib.call("mg_mlp_backward", b, _lib.ptr(x), _lib.ptr(tgt), _lib.ptr(self.params),
                  _lib.ptr(self.wbf16), _lib.ptr(work), _lib.ptr(lp), _lib.stream_handle())
        grad = torch.empty_like(self.params)
        # an Adam pass with lr = 0 sums the gradient slices without updating
        _lib.call("mg_mlp_adam", _lib.ptr(self.params), _lib.ptr(self.m), _lib.ptr(self.v),
                  _lib.ptr(work), b, 1, 0.0, 0.9, 0.999, 1e-8, _lib.ptr(self.wbf16),
                  _lib.ptr(grad), _lib.stream_handle())
        return float(lp.double().sum()), grad

    def _work(self, b):
        nbytes = _lib.load().mg_mlp_workspace_bytes(b)
        if nbytes < 0:
            raise ValueError("batch must be a positive multiple of 256")
        cache = getattr(self, "_work_cache", {})
        if b not in cache:
            cache[b] = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._work_cache = cache
        return cache[b]

    def train(self, x, tgt, steps, batch=8192, lr=2e-3, lr_min=2e-5, seed=0,
              log_every=250, beta1=0.9, beta2=0.999, eps=1e-8):
        """Adam on minibatches drawn with replacement from ``(x, tgt)``;
        cosine learning-rate decay from ``lr`` to ``lr_min``."""
        n = x.shape[0]
        work = self._work(batch)
        lp = torch.empty((batch * OUT_PAD + 255) // 256, dtype=torch.float32,
                         device=self.device)
        g = torch.Generator(device=self.device).manual_seed(seed)
        stream = _lib.stream_handle()
        lib = _lib.load()
        for it in range(steps):
            idx = torch.randint(n, (batch,), device=self.device, generator=g)
            xb = x.index_select(0, idx)
            tb = tgt.index_select(0, idx)
            _lib.check(lib.mg_mlp_forward(batch, _lib.ptr(xb), _lib.ptr(self.params),
                                          _lib.ptr(self.wbf16), _lib.ptr(work), None,
                                          stream))
            _lib.check(lib.mg_mlp_backward(batch, _lib.ptr(xb), _lib.ptr(tb),
                                           _lib.ptr(self.params), _lib.ptr(self.wbf16),
                                           _lib.ptr(work), _lib.ptr(lp), stream))
            self.step += 1
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * it / steps))
            _lib.check(lib.mg_mlp_adam(_lib.ptr(self.params), _lib.ptr(self.m),
                                       _lib.ptr(self.v), _lib.ptr(work), batch,
                                       self.step, lr_t, beta1, beta2, eps,
                                       _lib.ptr(self.wbf16), None, stream))
            if log_every and (it % log_every == 0 or it == steps - 1):
                self.history.append((self.step, float(lp.double().sum())))
        return self.history

    def predict_dy(self, states, batch=8192):
        """``dY [n, 7]`` (FP64) predicted for states ``[n, 10]``."""
        n = states.shape[0]
        x, _ = self.features(states)
        pad = (-n) % 256
        if pad:
            x = torch.cat([x, torch.zeros((pad, IN_PAD), dtype=x.dtype,
                                          device=x.device)])
        outs = []
        for s in range(0, x.shape[0], batch):
            xb = x[s:s + batch]
            if xb.shape[0] % 256:
                raise AssertionError("padded batch")
            outs.append(self.forward(xb))
        out = torch.cat(outs)[:n, :N_OUT].double()
        mean = torch.tensor(list(self.norm.out_mean), dtype=torch.float64,
                            device=out.device)
        scale = torch.tensor(list(self.norm.out_scale), dtype=torch.float64,
                             device=out.device)
        return out * scale + mean

    # }}}

    # {{{ inference in the right-hand side

    def species_source(self, cv, temperature, dt, nodes_per_element, rhs=None,
                       fallback_counter=None):
        """Surrogate ``W_k omega_k``: added into the species rows of ``rhs``
        (a ``ConservedVars``; returned) or, without ``rhs``, as a
        ``ConservedVars`` of sources (zero outside the species rows)."""
        from mirgecom_amd.fluid import ConservedVars
        _lib.ensure_mechanism()
        q = cv.q.contiguous()
        t = temperature.contiguous()
        n = cv.mass.numel()
        gas = _lib.MgGas(eos_kind=1, nspecies=7)
        if rhs is not None:
            r = rhs.q
            if r.shape != q.shape or r.dtype != torch.float64 or not r.is_contiguous():
                raise ValueError("rhs must be a contiguous float64 tensor shaped "
                                 "like the state")
            out, add, ret = r, 1, rhs
        else:
            full = torch.zeros_like(q)
            out, add = full[2 + cv.dim:], 0
            ret = ConservedVars(cv.dim, full)
        _lib.call("mg_chem_surrogate_source", gas, cv.dim, int(nodes_per_element), n,
                  _lib.ptr(q), _lib.ptr(t), float(dt), _lib.ptr(self.params),
                  _lib.ptr(self.wbf16), self.norm, add, _lib.ptr(out),
                  _lib.ptr(fallback_counter), _lib.stream_handle())
        return ret

    # }}}

    # {{{ persistence (plain arrays, nothing executable)

    def state_dict(self):
        nm = self.norm
        return {"params": self.params.cpu().numpy(),
                "norm": np.concatenate([np.array(nm.in_mean), np.array(nm.in_scale),
                                        np.array(nm.out_mean), np.array(nm.out_scale),
                                        np.array(nm.lo), np.array(nm.hi),
                                        [nm.check_tol, nm.sample_period,
                                         nm.sample_offset]])}

    def load_state_dict(self, d):
        self.params.copy_(torch.as_tensor(d["params"], device=self.device))
        a = np.asarray(d["norm"], dtype=np.float64)
        nm = self.norm
        o = 0
        for name, k in (("in_mean", 10), ("in_scale", 10), ("out_mean", 7),
                        ("out_scale", 7), ("lo", 10), ("hi", 10)):
            arr = getattr(nm, name)
            for i in range(k):
                arr[i] = a[o + i]
            o += k
        nm.check_tol = float(a[o])
        nm.sample_period = int(a[o + 1])
        nm.sample_offset = int(a[o + 2])
        self.sync_bf16()

    # }}}


__all__ = ["ChemistrySurrogate", "autoignition_dataset", "reactor_trajectories",
           "states_from_field", "substep_targets", "premixed_ethylene_air",
           "surrogate_reactor"]
