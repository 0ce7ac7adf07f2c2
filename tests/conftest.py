import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (runs the HIP kernels)")


def element_index(elements):
    lut = {e: i for i, e in enumerate(sorted(set(elements)))}
    return np.array([lut[e] for e in elements], dtype=np.int32)


@pytest.fixture(scope="session")
def small_random():
    """600 atoms, two species, 20 A cube (box coordinates)."""
    from fullrmc_amd import systems
    s = systems.random_box(600, 20.0, 1.8, seed=3, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    s["elem"] = element_index(s["elements"])
    s["box"] = systems.box_coordinates(s["recip"], s["real"])
    return s


@pytest.fixture(scope="session")
def small_triclinic():
    """40 rigid 13-atom rings (520 atoms, 3 species) in a skewed cell."""
    from fullrmc_amd import systems
    s = systems.ring_box(nmol=40, cell=(18.0, 19.0, 20.0, 80.0, 95.0, 105.0), seed=5, min_centre=4.0)
    s["elem"] = element_index(s["elements"])
    s["box"] = systems.box_coordinates(s["recip"], s["real"])
    return s
