"""Puncture configurations and grid parameters, mapped onto the reference paramstruct.

The shipped defaults are NRPyElliptic_Ccodes/free_parameters.h:1-41 (CFL 0.7, AMAX 1e6,
SINHWAA 0.07, bScale 5, eta 20). The workloads the benchmarks and tests use:
shipped_free_parameters, axisymmetric_spinning (64x64x4), equal_mass_boosted (128x128x16),
unequal_mass_q4 (seed-7 spins) and random_batch (seeded draws of q, D, momenta, spins and grids;
all draws from numpy's default_rng).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence

import numpy as np

from . import abi


@dataclasses.dataclass
class PunctureConfig:
    """Two Bowen-York punctures plus the SinhSymTP grid they are solved on."""
    Nxx: Sequence[int] = (128, 128, 16)
    bare_mass: Sequence[float] = (0.5, 0.5)
    pos0: Sequence[float] = (0.0, 0.0, 5.0)
    pos1: Sequence[float] = (0.0, 0.0, -5.0)
    P0: Sequence[float] = (0.0, 0.0, 0.0)
    P1: Sequence[float] = (0.0, 0.0, 0.0)
    S0: Sequence[float] = (0.0, 0.0, 0.0)
    S1: Sequence[float] = (0.0, 0.0, 0.0)
    AMAX: float = 1.0e6            # free_parameters.h:10
    SINHWAA: float = 0.07          # free_parameters.h:11
    bScale: float = 5.0            # free_parameters.h:13
    eta: float = 20.0              # free_parameters.h:21
    CFL: float = 0.7               # free_parameters.h:6
    name: str = ""

    def fill_params(self, p: abi.paramstruct) -> None:
        """Overwrite the physics members of a paramstruct (as free_parameters.h does)."""
        p.time = 0.0
        p.bScale, p.AMAX, p.SINHWAA = self.bScale, self.AMAX, self.SINHWAA
        p.eta_damping = self.eta
        p.bare_mass_0, p.bare_mass_1 = self.bare_mass
        p.puncture_0_x, p.puncture_0_y, p.puncture_0_z = self.pos0
        p.puncture_1_x, p.puncture_1_y, p.puncture_1_z = self.pos1
        p.P0_x, p.P0_y, p.P0_z = self.P0
        p.P1_x, p.P1_y, p.P1_z = self.P1
        p.S0_x, p.S0_y, p.S0_z = self.S0
        p.S1_x, p.S1_y, p.S1_z = self.S1
        p.outer_bc_type = abi.RADIATION_OUTER_BCS

    def features(self) -> np.ndarray:
        """Per-configuration conditioning vector for the learned initial guess:
        q, D, P0, P1, S0, S1 (16 numbers)."""
        m0, m1 = self.bare_mass
        q = max(m0, m1) / max(min(m0, m1), 1e-12)
        D = float(np.linalg.norm(np.subtract(self.pos0, self.pos1)))
        return np.array([q, D, *self.P0, *self.P1, *self.S0, *self.S1, m0, m1], dtype=np.float64)


def shipped_free_parameters(Nxx=(128, 128, 16)) -> PunctureConfig:
    """The punctures hard-coded in NRPyElliptic_Ccodes/free_parameters.h:24-40."""
    return PunctureConfig(Nxx=tuple(Nxx), bare_mass=(0.5184199353358704, 0.39193567996522616),
                          pos0=(0.0, 0.0, 5.0), pos1=(0.0, 0.0, -5.0),
                          P0=(0.09530152296974252, 0.0, -0.00084541526517121),
                          P1=(-0.09530152296974252, 0.0, 0.00084541526517121),
                          S0=(0.0, 0.09509112426035504, 0.0), S1=(0.0, -0.09156449704142013, 0.0),
                          name="free_parameters_h")


def axisymmetric_spinning(Nxx=(64, 64, 4)) -> PunctureConfig:
    """64x64x4 CPU-runnable case: m=0.5 each at z=+-5, P=0, S=(0,0,0.2) on each."""
    return PunctureConfig(Nxx=tuple(Nxx), bare_mass=(0.5, 0.5), S0=(0.0, 0.0, 0.2), S1=(0.0, 0.0, 0.2),
                          name="axisym_spin_64x64x4")


def equal_mass_boosted(Nxx=(128, 128, 16)) -> PunctureConfig:
    """128x128x16 production case: m=0.5 each at z=+-5, P=(0,+-0.0953,0), no spins."""
    return PunctureConfig(Nxx=tuple(Nxx), bare_mass=(0.5, 0.5), P0=(0.0, 0.0953, 0.0),
                          P1=(0.0, -0.0953, 0.0), name="equal_mass_boosted_128x128x16")


def _random_unit(rng) -> np.ndarray:
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def binary_from_qD(q: float, D: float, p_scale: float, chi0: np.ndarray, chi1: np.ndarray,
                   Nxx, name="") -> PunctureConfig:
    """Unequal-mass binary on the z axis at z = +-D/2, total bare mass 1, Newtonian circular
    tangential momenta |P| = m0 m1 / sqrt(M D) (times p_scale) along +-y, spins S = chi m^2.
    The SinhSymTP foci are put on the punctures: bScale = D/2."""
    M = 1.0
    m0 = M * q / (1.0 + q)
    m1 = M - m0
    P = p_scale * m0 * m1 / math.sqrt(M * D)
    return PunctureConfig(Nxx=tuple(int(n) for n in Nxx), bare_mass=(m0, m1), pos0=(0.0, 0.0, D / 2),
                          pos1=(0.0, 0.0, -D / 2), P0=(0.0, P, 0.0), P1=(0.0, -P, 0.0),
                          S0=tuple(float(x) for x in chi0 * m0 * m0), S1=tuple(float(x) for x in chi1 * m1 * m1),
                          bScale=D / 2, name=name)


def unequal_mass_q4(Nxx=(128, 128, 16)) -> PunctureConfig:
    """q=4 (m=0.8/0.2) at D=10 on z, Newtonian-circular momenta, spins from seed 7 with
    isotropic directions and |chi| <= 0.6."""
    rng = np.random.default_rng(7)
    chi0 = _random_unit(rng) * rng.uniform(0.0, 0.6)
    chi1 = _random_unit(rng) * rng.uniform(0.0, 0.6)
    return binary_from_qD(4.0, 10.0, 1.0, chi0, chi1, Nxx, name="q4_D10_seed7_128x128x16")


def random_batch(n: int, seed: int, grids: bool = True, Nxx=(128, 128, 16)) -> List[PunctureConfig]:
    """`n` configurations: q log-uniform in [1,8], M=1, D~U(6,20), |P| = Newtonian x U(0.9,1.1),
    spins isotropic with |chi|~U(0,0.8); grids mixed from Nxx0 in {96,128,160},
    Nxx1 in {96,128}, Nxx2 in {8,16} (or all `Nxx` when grids=False)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        q = math.exp(rng.uniform(0.0, math.log(8.0)))
        D = rng.uniform(6.0, 20.0)
        ps = rng.uniform(0.9, 1.1)
        chi0 = _random_unit(rng) * rng.uniform(0.0, 0.8)
        chi1 = _random_unit(rng) * rng.uniform(0.0, 0.8)
        if grids:
            g = (int(rng.choice([96, 128, 160])), int(rng.choice([96, 128])), int(rng.choice([8, 16])))
        else:
            g = tuple(Nxx)
        out.append(binary_from_qD(q, D, ps, chi0, chi1, g, name=f"seed{seed}_{k}"))
    return out
