"""Fixed 3-D pulsed dielectric channel problem in normalized c = mu = 1 units
(drop-in for warp_fdtd/guide_problem.py:13-78; the octree ``mesh`` builder
belongs to the adaptive-mesh prototype and is not part of this package).

Lengths are micrometers; one time unit is one micrometer/c.  The matched E/H
conductivity layer is a sponge, not CPML.
"""

from dataclasses import asdict, dataclass

import numpy as np


@dataclass(frozen=True)
class GuideProblem:
    length: float = 4.8
    width: float = 3.2
    height: float = 3.2
    core_width: float = 0.4
    core_height: float = 0.4
    core_epsilon: float = 12.25
    clad_epsilon: float = 2.25
    wavelength: float = 1.55
    source_x: float = 0.8
    source_sigma_x: float = 0.10
    source_sigma_transverse: float = 0.12
    pulse_center: float = 3.0
    pulse_sigma: float = 1.0
    duration: float = 14.0
    sponge_width: float = 0.4
    sponge_rate: float = 10.0

    @property
    def lengths(self):
        return (self.length, self.width, self.height)

    def waveform(self, t):
        """guide_problem.py:72-75"""
        return np.exp(-0.5 * ((t - self.pulse_center) / self.pulse_sigma) ** 2) * np.sin(
            2 * np.pi * t / self.wavelength)

    def parameters(self):
        return asdict(self)
