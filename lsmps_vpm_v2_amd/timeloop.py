"""The reference's iteration loop restricted to the velocity path, on the GPU.

The reference's loop (main.cpp:296-429) runs, per step, penalization,
advection (main.cpp:373), diffusion, remeshing and the velocity solve
(main.cpp:410).  Penalization, diffusion and remeshing lie outside this
package; what remains is an inviscid vortex-particle loop: move the particles
with the current velocity, then solve for the velocity at the new positions.
The particle state stays resident on the GPU between steps, and each step is
one library call (``vpmfmm_step_device``: the advection kernel, then the FMM
solve, on one stream).

``advection`` mirrors the reference class of that name
(src/advection/advection.hpp:14-24, main_advection.cpp:6-83).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .solver import FmmSolver, Particle


@dataclass
class DeviceParticle:
    """Particle fields of the velocity path as GPU tensors (float64, uint8)."""
    x: "torch.Tensor"  # noqa: F821
    y: "torch.Tensor"  # noqa: F821
    s: "torch.Tensor"  # noqa: F821
    vorticity: "torch.Tensor"  # noqa: F821
    isActive: "torch.Tensor"  # noqa: F821
    u: "torch.Tensor"  # noqa: F821
    v: "torch.Tensor"  # noqa: F821

    @property
    def num(self) -> int:
        return self.x.numel()

    @classmethod
    def from_particle(cls, p: Particle, device="cuda") -> "DeviceParticle":
        import torch

        def d(a, dt=torch.float64):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(device)
        n = p.num
        u = p.u if len(p.u) == n else np.zeros(n)
        v = p.v if len(p.v) == n else np.zeros(n)
        return cls(d(p.x), d(p.y), d(p.s), d(p.vorticity), d(p.isActive, torch.uint8),
                   d(u), d(v))

    def to_particle(self) -> Particle:
        h = lambda t: t.detach().cpu().numpy().copy()  # noqa: E731
        return Particle(x=h(self.x), y=h(self.y), s=h(self.s), vorticity=h(self.vorticity),
                        isActive=h(self.isActive), u=h(self.u), v=h(self.v))


class advection:  # noqa: N801 -- reference class name
    """Advection sub-step on the GPU (main_advection.cpp:6-83).

    ``scheme`` "euler" is the reference's compiled choice (ADVECTION_TYPE 1);
    "rk2" is its alternative, which scales the current velocity only."""

    def __init__(self, solver: FmmSolver, dt: float, scheme: str = "euler"):
        self.solver, self.dt, self.scheme = solver, float(dt), scheme

    def main_advection(self, p: DeviceParticle) -> None:
        self.solver.advect_device(p.x, p.y, p.u, p.v, self.dt, self.scheme)

    def advection_euler(self, p: DeviceParticle) -> None:
        self.solver.advect_device(p.x, p.y, p.u, p.v, self.dt, "euler")

    def advection_rk2(self, p: DeviceParticle) -> None:
        self.solver.advect_device(p.x, p.y, p.u, p.v, self.dt, "rk2")


class InviscidLoop:
    """Advection + velocity, one library call per step, state on the GPU."""

    def __init__(self, particle: DeviceParticle, dt: float, scheme: str = "euler",
                 solver: FmmSolver | None = None, **solver_opts):
        self.p = particle
        self.dt = float(dt)
        self.scheme = scheme
        self.solver = solver or FmmSolver(**solver_opts)
        self.step_count = 0

    def start(self) -> None:
        """Velocity at the initial positions.  (In the reference it comes with
        the initial state: the pre-loop solve at main.cpp:236 is commented
        out, and the loop's own solve closes each step, main.cpp:410.)"""
        p = self.p
        self.solver.velocity_device(p.x, p.y, p.s, p.vorticity, p.isActive, p.u, p.v)

    def step(self) -> None:
        p = self.p
        self.solver.step_device(p.x, p.y, p.s, p.vorticity, p.isActive, p.u, p.v, self.dt,
                                self.scheme)
        self.step_count += 1

    def run(self, steps: int, callback=None) -> None:
        for _ in range(steps):
            self.step()
            if callback is not None:
                callback(self)
