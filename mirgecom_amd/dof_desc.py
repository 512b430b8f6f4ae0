"""Domain descriptors of the call surface (the subset of grudge's
``grudge.dof_desc`` the reference's operators use: mirgecom/euler.py:367-372,
mirgecom/navierstokes.py:448-451)."""

from __future__ import annotations

from dataclasses import dataclass, replace

DISCR_TAG_BASE = "base"
DISCR_TAG_QUAD = "quad"
BTAG_ALL = "all"
FACE_RESTR_ALL = "all_faces"


@dataclass(frozen=True)
class VolumeDomainTag:
    tag: object = None


@dataclass(frozen=True)
class BoundaryDomainTag:
    tag: object
    volume_tag: object = None


@dataclass(frozen=True)
class DOFDesc:
    domain_tag: object
    discretization_tag: object = DISCR_TAG_BASE

    def with_discr_tag(self, tag):
        return replace(self, discretization_tag=tag)

    def with_domain_tag(self, dtag):
        if not isinstance(dtag, (VolumeDomainTag, BoundaryDomainTag)):
            dtag = BoundaryDomainTag(dtag)
        return replace(self, domain_tag=dtag)


DD_VOLUME_ALL = DOFDesc(VolumeDomainTag(), DISCR_TAG_BASE)


def as_dofdesc(x):
    if isinstance(x, DOFDesc):
        return x
    if x in ("vol", None):
        return DD_VOLUME_ALL
    return DOFDesc(BoundaryDomainTag(x), DISCR_TAG_BASE)
