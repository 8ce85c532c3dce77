"""The pluggable Poisson-solver interface (reference solver/base.py:27-41).

A solver maps mass density ``rho`` to the potential ``phi`` solving lap(phi) = 4*pi*G*rho
with open (isolated) boundaries on the node-centred grid; ``solve`` reads ``rho`` and
overwrites ``phi`` in place. ``warm_start=True`` lets iterative solvers start from the
current contents of ``phi``.
"""

from __future__ import annotations

from abc import ABC, abstractmethod


class PoissonSolver(ABC):
    grid_size: int
    dx: float

    @abstractmethod
    def solve(self, rho, phi, warm_start: bool = False) -> None:
        raise NotImplementedError


__all__ = ["PoissonSolver"]
