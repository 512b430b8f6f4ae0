"""Call-surface helpers (mirgecom/utils.py:139-148)."""

from __future__ import annotations

from mirgecom_amd.dof_desc import BoundaryDomainTag


def normalize_boundaries(boundaries):
    """Key a boundaries dict by ``BoundaryDomainTag`` (mirgecom/utils.py:
    139-148)."""
    return {
        (btag if isinstance(btag, BoundaryDomainTag)
         else BoundaryDomainTag(btag)): bdry
        for btag, bdry in boundaries.items()}
