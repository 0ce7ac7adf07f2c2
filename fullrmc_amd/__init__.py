"""fullrmc_amd -- fullrmc's pair-distance / pair-histogram hot path on AMD
MI355X (gfx950).

* ``fullrmc_amd.Core.{pairs_distances, pairs_histograms,
  boundary_conditions_collection}``: drop-ins for the reference's Cython
  extensions, same signatures, computed by hand-written HIP kernels in
  ``libfrmc_hip.so`` (C ABI in ``include/frmc.h``).
* ``fullrmc_amd.Engine`` / ``fullrmc_amd.Constraints``: the run loop and the
  PairDistribution / StructureFactor constraints with the reference's
  semantics, either calling the drop-ins per move (``Engine.run``) or
  running device-resident with speculative k-proposal windows
  (``Engine.run_device``).
"""
__version__ = "0.1.0"
