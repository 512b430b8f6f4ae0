"""HIP-graph capture of right-hand sides and time steps.

A device RHS is two to four kernel launches plus a few hundred microseconds of
Python dispatch; at config sizes the kernels themselves take ~0.3 ms, so the
host is a visible share of a step.  Capturing the launches once into a HIP
graph (``torch.cuda.CUDAGraph``, which records the library's launches on the
capture stream) and replaying it removes the per-call host work.  This is the
device path's counterpart of the reference's ``actx.compile(rhs)``
(mirgecom/steppers.py:49-59).

Constraints of a captured call: the state shape, the gas model, the
boundaries and every Python scalar the RHS uses (time, dt) are fixed at
capture; boundary callbacks must not synchronise with the host.  The input
and output buffers are owned by the graph object.
"""

from __future__ import annotations

import torch

from mirgecom_amd.fluid import ConservedVars


def _warm(fn, n):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(n):
            fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()


class GraphedRHS:
    """``rhs(t, cv)`` captured at a fixed ``t`` for states shaped like
    ``cv``.  ``graphed.state`` is the input buffer (write the state there, or
    call the object with a state, which copies it in); ``graphed.replay()``
    evaluates into ``graphed.out`` without any host work besides the launch."""

    def __init__(self, rhs, cv, t=0.0, warmup=2):
        self.dim = cv.dim
        self.state = cv.q.detach().clone().contiguous()
        arg = ConservedVars(self.dim, self.state)
        _warm(lambda: rhs(t, arg), warmup)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = rhs(t, arg).q

    def replay(self):
        self.graph.replay()
        return ConservedVars(self.dim, self.out)

    def __call__(self, t, cv):
        self.state.copy_(cv.q)
        return self.replay()


class GraphedStepper:
    """One ``timestepper(state, t, dt, rhs)`` step captured twice, from
    buffer A to B and from B to A, so that stepping replays alternately with
    no copies.  ``t`` and ``dt`` are fixed at capture (time-independent RHS)."""

    def __init__(self, timestepper, rhs, cv, dt, t=0.0, warmup=1):
        self.dim = cv.dim
        self.buf = [cv.q.detach().clone().contiguous(),
                    torch.empty_like(cv.q).contiguous()]
        self.graphs = []
        for src, dst in ((0, 1), (1, 0)):
            arg = ConservedVars(self.dim, self.buf[src])

            def step(arg=arg, dst=dst):
                out = timestepper(state=arg, t=t, dt=dt, rhs=rhs)
                self.buf[dst].copy_(out.q)
            _warm(step, warmup)
            self.buf[0].copy_(cv.q)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                step()
            self.graphs.append(g)
        self.buf[0].copy_(cv.q)
        self.cur = 0

    def step(self):
        """Advance the held state by one step; returns it."""
        self.graphs[self.cur].replay()
        self.cur ^= 1
        return ConservedVars(self.dim, self.buf[self.cur])

    @property
    def state(self):
        return ConservedVars(self.dim, self.buf[self.cur])
