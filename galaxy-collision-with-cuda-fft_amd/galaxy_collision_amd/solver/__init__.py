"""Poisson solvers behind ``make_solver`` (reference solver/__init__.py:12-22).

Names:
- ``multigrid``: open-BC multigrid, float32, the reference's algorithm and defaults;
- ``multigrid_mp``: the same discrete problem solved by float64 defect correction around
  float32 V-cycles (mixed precision);
- ``fft``: zero-padded float64 FFT convolution with the continuum Green's function on the
  device (the reference's validation solver);
- ``multigrid_learned``: ``multigrid_mp`` warm-started by time extrapolation plus the
  learned correction (``net=WarmStartNet`` or ``weights=path``; by default the packaged
  weights for grid_size/pool, if present).
"""

from galaxy_collision_amd.solver.base import PoissonSolver

SOLVER_NAMES = ("multigrid", "multigrid_mp", "fft", "multigrid_learned")


def make_solver(name: str, grid_size: int, **kwargs) -> PoissonSolver:
    """Construct a solver by name (see the module docstring)."""
    if name == "multigrid":
        from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

        return MultigridPoissonSolver(grid_size, **kwargs)
    if name == "multigrid_mp":
        from galaxy_collision_amd.solver.multigrid import MixedPrecisionMultigridSolver

        return MixedPrecisionMultigridSolver(grid_size, **kwargs)
    if name == "multigrid_learned":
        from galaxy_collision_amd.solver.multigrid import MixedPrecisionMultigridSolver
        from galaxy_collision_amd.warmstart import WEIGHTS_DIR, WarmStartNet

        net = kwargs.pop("net", None)
        weights = kwargs.pop("weights", None)
        pool = int(kwargs.pop("pool", 4))
        if net is None:
            path = weights or WEIGHTS_DIR / f"wsnet_r{grid_size // pool}.npz"
            net = WarmStartNet.load(path)
        return MixedPrecisionMultigridSolver(grid_size, warm="learned", net=net, pool=pool,
                                             **kwargs)
    if name == "fft":
        from galaxy_collision_amd.solver.fft import FFTPoissonSolver

        return FFTPoissonSolver(grid_size, **kwargs)
    raise ValueError(f"unknown solver {name!r} (expected one of {SOLVER_NAMES})")


__all__ = ["PoissonSolver", "make_solver", "SOLVER_NAMES"]
