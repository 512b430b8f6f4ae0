"""Low-level numerical fluxes (mirgecom/flux.py:45-177), pointwise on tensors.

The operators evaluate these inside the HIP kernels; the tensor versions
serve direct callers and identify the flux choice of an operator call."""

from __future__ import annotations

import torch


def num_flux_lfr(f_minus_normal, f_plus_normal, q_minus, q_plus, lam):
    """mirgecom/flux.py:45-84"""
    return (f_minus_normal + f_plus_normal + lam * (q_minus - q_plus)) / 2


def num_flux_central(f_minus_normal, f_plus_normal):
    """mirgecom/flux.py:87-110"""
    return (f_plus_normal + f_minus_normal) / 2


def num_flux_hll(f_minus_normal, f_plus_normal, q_minus, q_plus, s_minus,
                 s_plus):
    """mirgecom/flux.py:113-177"""
    f_star = (s_plus * f_minus_normal - s_minus * f_plus_normal
              + s_plus * s_minus * (q_plus - q_minus)) / (s_plus - s_minus)
    f = torch.where(s_minus >= 0, f_minus_normal, f_star)
    return torch.where(s_plus <= 0, f_plus_normal, f)
