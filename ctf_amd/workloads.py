"""The turbulence set-ups used for parity tests and measurements.

Each workload is the unit cube, periodic, gamma-law with Gamma = 4/3, uniform rho = 1,
p = 1 and a uniform field along z, plus seeded random-phase Fourier perturbations
(see ``csrc/initial_data.cpp`` for the definitions):

* ``c1``  64^3, v rms 0.3 on 1 <= |k| <= 4, Bz = 1; PLM (MC, theta = 2) / HLLD / RK2,
          CFL 0.4, FP64, 10 steps (the reference's CPU-runnable case)
* ``c2``  256^3, 4-velocity u = W v of rms 0.5 on |k| <= 8 (a Gaussian 3-velocity of rms
          0.5 exceeds c at its peaks), dB rms 0.5 from a vector potential on |k| <= 4
          (discretely solenoidal), Bz = 1; WENO5 / HLLD / RK3, 50 steps, FP64
* ``c3``  c2 in the mixed form (FP32 reconstruction and Riemann fluxes)
* ``c4``  c3 with the learned c2p initial guess
* ``c5``  512^3, 4-velocity rms 0.9 on |k| <= 8, Bz = 3; WENO5 / HLLD / RK3, 20 steps,
          mixed + learned
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace


@dataclass(frozen=True)
class Workload:
    name: str
    N: int
    extrap: str
    riemann: str
    advance: str
    steps: int
    cfl: float = 0.4
    precision: str = "fp64"
    c2p: str = "plain"
    v_rms: float = 0.3
    v_kmax: float = 4.0
    four_velocity: bool = False
    Bz: float = 1.0
    dB_rms: float = 0.0
    A_kmax: float = 4.0
    seed: int = 20240601
    gamma: float = 4.0 / 3.0
    Ng: int = 3
    godunov: str = "plm-split"
    eigen: str = "component"

    def ic_kwargs(self, N=None):
        n = N or self.N
        return dict(N=(n, n, n), Ng=self.Ng, rho=1.0, pre=1.0, B0=(0.0, 0.0, self.Bz),
                    v_rms=self.v_rms, v_kmin=1.0, v_kmax=self.v_kmax,
                    four_velocity=self.four_velocity, dB_rms=self.dB_rms, A_kmin=1.0,
                    A_kmax=self.A_kmax, seed=self.seed)

    def config_kwargs(self, N=None):
        n = N or self.N
        return dict(N=(n, n, n), Ng=self.Ng, gamma=self.gamma, extrap=self.extrap, theta=2.0,
                    IS="js96", riemann=self.riemann, advance=self.advance,
                    precision=self.precision, c2p=self.c2p, godunov=self.godunov,
                    eigen=self.eigen)

    def resized(self, N):
        return replace(self, N=N)


C1 = Workload("c1", 64, "plm", "hlld", "rk2", 10, v_rms=0.3, v_kmax=4.0, Bz=1.0)
C2 = Workload("c2", 256, "weno5", "hlld", "rk3", 50, v_rms=0.5, v_kmax=8.0, four_velocity=True,
              Bz=1.0, dB_rms=0.5, A_kmax=4.0)
C3 = replace(C2, name="c3", precision="mixed")
C4 = replace(C3, name="c4", c2p="learned")
C5 = Workload("c5", 512, "weno5", "hlld", "rk3", 20, precision="mixed", c2p="learned",
              v_rms=0.9, v_kmax=8.0, four_velocity=True, Bz=3.0)

ALL = {w.name: w for w in (C1, C2, C3, C4, C5)}
