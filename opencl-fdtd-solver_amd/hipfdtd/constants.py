"""SI electromagnetic constants (same values as the reference's constants.py:11-14)."""

from __future__ import annotations

import math

C0 = 299_792_458.0
MU0 = 4e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0**2)
ETA0 = math.sqrt(MU0 / EPS0)

__all__ = ["C0", "MU0", "EPS0", "ETA0"]
