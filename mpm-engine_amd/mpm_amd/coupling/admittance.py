"""First-order force admittance for a vertical tool (src/warpmpm/coupling/admittance.py:
14-43): the commanded descent speed falls linearly with the measured reaction and reaches
zero at the target force (reverses above it when retraction is allowed)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class ForceAdmittance:
    f_target: float = 40.0
    v_max: float = 0.4
    damping: float | None = None  # N / (m/s); default f_target / v_max
    allow_retract: bool = True

    def _damp(self) -> float:
        return self.damping if self.damping is not None else self.f_target / self.v_max

    def velocity_down(self, f_react: float) -> float:
        """Downward speed (m/s, positive descending) for an upward reaction f_react (N)."""
        v = (self.f_target - float(f_react)) / self._damp()
        return float(np.clip(v, -self.v_max if self.allow_retract else 0.0, self.v_max))
