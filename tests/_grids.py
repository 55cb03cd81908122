"""Real grids for the tests: the run-phase files the reference's GetGrid
wrote for two of its examples (tests/golden/grids, made by
tests/golden/make_grids.py) -- Monterey Bay (examples/tides: 1534 cells,
stair-stepped bathymetry on 10 levels, velocity and specified-value open
boundaries) and the internal-wave tank (examples/iwaves: 200 cells on 100
stretched levels) -- with seeded state as cases.from_files sets it, and Monterey Bay generated
from the example's own planform."""
import os

from suntans_amd import cases

HERE = os.path.dirname(os.path.abspath(__file__))


def path(name: str) -> str:
    return os.path.join(HERE, "golden", "grids", name)


REAL = [("mbay", lambda: cases.from_files(path("tides"), seed=3)),
        ("iwaves", lambda: cases.from_files(path("iwaves"), seed=3)),
        # Monterey Bay generated here from the example's planform
        # (sun_grid_generate): the grid GetGrid leaves in memory
        ("mbay_gen", lambda: cases.from_planform(os.path.join(path("tides"), "planform"), seed=5))]
