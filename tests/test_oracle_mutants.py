"""Each pin is a real check: the oracle built with a planted mistake
(ORACLE_MUTANT k in oracle/wig_oracle.c, `make -C oracle mutants`) fails the
pin that covers that part of the step, while the oracle itself passes it
(tests/test_oracle_pins.py)."""
import os
import subprocess

import pytest

from oracle import oracle
from tests import oracle_pins

ORACLE_DIR = os.path.dirname(os.path.abspath(oracle.__file__))

# planted mistake -> the pin that must catch it
CAUGHT_BY = {
    1: "advect",                   # AFFINE_KEEP 1.0 instead of 0.95
    2: "prolong",                  # prolongation weights 1:1:1:1 instead of 9:3:3:1
    3: "restrict",                 # a coarse cell with an air child is not air
    4: "restrict",                 # the coarse rhs the mean of the fluid children's residuals, not a quarter of their sum
    5: "smooth",                   # coarse smoothing without the level's h^2
    6: "coarsest",                 # the coarsest SOR at omega 1 instead of 1.8
    7: "advect",                   # speed limit 16 cells a step instead of 8
    8: "advect",                   # a wall does not zero the velocity component
    9: "sort",                     # G_START an inclusive prefix sum
    10: "raster",                  # a majority tie goes to the higher type
    11: "raster",                  # the cell word from the highest id
    12: "extrapolate",             # the second layer from projected faces only
    13: "project",                 # the projection's sign
    14: "classify",                # an empty cell with one fluid neighbour counted fluid
    15: "classify",                # no gap ever sealed
    16: "sync_particles",          # coincident particles not pushed apart
    17: "sync_particles_shuffled", # rankById by slot order, not by id
    18: "raster",                  # the mean velocity a float sum, not an exact one
    19: "relax",                   # the fine smoother without the ghost fluid's weights toward air
    20: "compact",                 # a block worked on only if it is marked itself
    21: "set_faces",               # gravity on the u faces
    22: "escape",                  # a buried particle stays buried
    23: "unpaint",                 # a cell taken by displaced air unpainted anyway
    24: "divergence",              # the drift rhs's sign
    25: "air_push",                # the air's pressure pushes up its gradient
    26: "surface_drag",            # the drag's push on the air without the mass ratio
    27: "push_gas_out",            # displaced gas's vapour lost
    28: "occupants",               # the person's y velocity decoded from the x field
    29: "occupants",               # liquid left drawn under a person
    30: "gravity",                 # the field's bilinear weights mirrored in x
    31: "gravity",                 # down taken as Earth's whatever the field
    32: "viscosity",               # the sweeps unrelaxed (omega 1)
    33: "viscosity",               # an obstacle's face across the flow counted once
    34: "surface",                 # the curvature's sign flipped
    35: "give_way",                # a push's turn over the body's mass, not its moment of inertia
    36: "body_face",               # the face's arm r x n taken n x r
}


@pytest.fixture(scope="module", autouse=True)
def mutants():
    r = subprocess.run(["make", "-s", "-j8", "-C", ORACLE_DIR, "mutants"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("k", sorted(CAUGHT_BY))
def test_planted_mistake_is_caught(k):
    lib = oracle.lib(oracle.mutant_path(k))
    with pytest.raises(AssertionError):
        oracle_pins.PINS[CAUGHT_BY[k]](lib)
