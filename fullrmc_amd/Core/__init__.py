"""gfx950 implementations of fullrmc's hot-path extension modules:
``pairs_distances``, ``pairs_histograms`` and ``boundary_conditions_collection``
(Extensions/*.pyx in the reference), same function signatures."""
