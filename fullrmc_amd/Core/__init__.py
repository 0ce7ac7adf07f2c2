"""gfx950 implementations of fullrmc's hot-path extension modules:
``pairs_distances``, ``pairs_histograms``, ``boundary_conditions_collection``
and ``reciprocal_space`` (Extensions/*.pyx in the reference), same function
signatures."""
