"""Time marching (call surface of mirgecom/steppers.py:76-195, 350-449).

``advance_state`` drives a stepper function ``timestepper(state, t, dt, rhs)``
(for example ``mirgecom_amd.integrators.rk4_step``) with the reference's pre-
and post-step callbacks, ``max_steps`` and local-dt modes.  There is no
tracing compiler on this path: each RHS is already a pair of fused HIP
kernels, so ``compile_rhs`` and ``force_eval`` are accepted and have no
effect.  Leap ``MethodBuilder`` steppers are not supported."""

from __future__ import annotations

import numpy as np


def advance_state(rhs, timestepper, state, t_final, t=0, istep=0, dt=0,
                  max_steps=None, component_id="state", pre_step_callback=None,
                  post_step_callback=None, force_eval=None, local_dt=False,
                  compile_rhs=True):
    """Advance ``state`` from ``t`` to ``t_final`` (or, with ``local_dt``,
    from step ``istep`` to ``max_steps``); returns ``(istep, t, state)``.

    Callbacks have the signature ``state, dt = callback(state=, step=, t=,
    dt=)``; the pre-step one runs before each step, the post-step one after
    the step counter and time have advanced (mirgecom/steppers.py:134-193)."""
    if type(timestepper).__name__ == "MethodBuilder":
        raise NotImplementedError("leap method builders are not supported")
    if local_dt:
        if max_steps is None:
            raise ValueError("max_steps must be given for local_dt mode.")
        marching_loc, marching_limit = istep, max_steps
    else:
        t = np.float64(t)
        marching_loc, marching_limit = t, t_final

    while marching_loc < marching_limit:
        if max_steps is not None and max_steps <= istep:
            return istep, t, state
        if pre_step_callback is not None:
            state, dt = pre_step_callback(state=state, step=istep, t=t, dt=dt)
        state = timestepper(state=state, t=t, dt=dt, rhs=rhs)
        istep += 1
        t = t + dt
        marching_loc = istep if local_dt else t
        if post_step_callback is not None:
            state, dt = post_step_callback(state=state, step=istep, t=t, dt=dt)
    return istep, t, state
