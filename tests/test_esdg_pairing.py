"""Host-side pieces of the paired flux-differencing loop of
``mg_euler_rhs_esdg`` (3-D, p = 3: 64 volume points, 192 face points):
the operators along cyclic diagonals are a re-indexing of the hybridized
operators, and the kernel's schedule -- volume pairs (h, h+s mod 64) for
s = 1..31 with the partner rows' sums moving one lane per step, s = 32 on
both lanes, face rows against the volume points (l+s) mod 64 -- visits every
nonzero entry of (Q_h - Q_h^T) o F_S once per row, so that it equals the
row-by-row sum for any symmetric two-point function."""

import numpy as np
import pytest

from mirgecom_amd.refelem import get_hybrid_operators


@pytest.mark.parametrize("dim,order", [(2, 2), (3, 3)])
def test_diagonal_layouts(dim, order):
    ho = get_hybrid_operators(dim, order, 2 * order)
    nq = ho.skew.shape[1]
    sd, ed = ho.skew_diag, ho.extrap_diag
    back = np.zeros_like(ho.skew)
    for s in range(nq):
        for h in range(nq):
            back[:, h, (h + s) % nq] = sd[:, s, h]
    assert np.array_equal(back, ho.skew)
    nft = ho.extrap.shape[0]
    back = np.zeros_like(ho.extrap)
    for s in range(nq):
        for m in range(nft):
            back[m, (m + s) % nq] = ed[s, m]
    assert np.array_equal(back, ho.extrap)


def test_paired_schedule_matches_rows():
    ho = get_hybrid_operators(3, 3, 6)
    nq, nft = 64, 192
    assert ho.skew.shape[1] == nq and ho.extrap.shape[0] == nft
    rng = np.random.default_rng(3)
    rx = rng.standard_normal((3, 3))
    bn = rng.standard_normal((nft, 3))   # fscale * w * n per face point
    g = rng.standard_normal((nq + nft, nq + nft, 3))
    g = g + g.transpose(1, 0, 2)         # symmetric two-point "flux"
    # row by row
    ref = np.zeros(nq + nft)
    for h in range(nq):
        ref[h] = sum(g[h, j] @ (rx.T @ ho.skew[:, h, j]) for j in range(nq))
        ref[h] += sum(g[h, nq + m] @ (ho.extrap[m, h] * bn[m]) for m in range(nft))
    for m in range(nft):
        ref[nq + m] = -sum(g[nq + m, j] @ (ho.extrap[m, j] * bn[m]) for j in range(nq))
    # the kernel's schedule; np.roll(v, -1)[l] = v[l + 1] is the one-lane move
    lane = np.arange(nq)
    sd, ed = ho.skew_diag, ho.extrap_diag
    acc, mv = np.zeros(nq), np.zeros(nq)
    for s in range(1, 33):
        fv = np.array([g[h, (h + s) % nq] @ (rx.T @ sd[:, s, h]) for h in lane])
        acc += fv
        if s < 32:
            mv = np.roll(mv - fv, -1)
    vol = acc + np.roll(mv, -32)
    face = np.zeros(nft)
    for w in range(3):
        m = w * nq + lane
        acc, mv = np.zeros(nq), np.zeros(nq)
        for s in range(nq):
            fv = np.array([g[nq + mm, (l + s) % nq] @ (ed[s, mm] * bn[mm])
                           for l, mm in zip(lane, m)])
            acc -= fv
            mv = np.roll(mv + fv, -1)
        vol += mv
        face[m] = acc
    got = np.concatenate([vol, face])
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()
