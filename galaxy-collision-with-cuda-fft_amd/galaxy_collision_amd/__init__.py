"""galaxy_collision_amd — an MI355X-native particle-mesh force chain.

Drop-in counterparts of the reference galaxy_collision package's hot path: ``deposit``
(CIC deposit / gradient / gather), ``integrator`` (kick, drift, kdk_step), ``solver``
(make_solver / PoissonSolver with open-boundary multigrid), ``diagnostics_device`` and the
``chain`` fused PM step. The compute runs in hand-written HIP kernels in libpmchain.so;
torch provides device memory and streams.
"""

__version__ = "0.1.0"
