"""Centre-of-mass container used by sp_properties (moment.py:4-11)."""
from dataclasses import dataclass, field

import torch


@dataclass
class Moment:
    cmtot: torch.Tensor = field(default_factory=lambda: torch.zeros(3, dtype=torch.float64))


def init_moment(cmtot):
    return Moment(torch.as_tensor(cmtot, dtype=torch.float64).clone())
