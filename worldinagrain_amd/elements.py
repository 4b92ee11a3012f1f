"""The element ids and kinds the liquid step reads (js/elements.js:8,416-457).

Ids are the element table's order: AIR 0, WALL 1, SAND 2, WATER 3, STONE 4,
GRAVEL 5.  `kinds()` is the `elems[t].kind` column as the C ABI takes it."""
import numpy as np

KIND_SOLID, KIND_POWDER, KIND_LIQUID, KIND_GAS = 1, 2, 3, 4
AIR, WALL, SAND, WATER, STONE, GRAVEL = 0, 1, 2, 3, 4, 5
FL_DRAWN = 4

_KINDS = {AIR: KIND_GAS, WALL: KIND_SOLID, SAND: KIND_POWDER, WATER: KIND_LIQUID, STONE: KIND_SOLID,
          GRAVEL: KIND_POWDER}

# Water's weight, Earth's g in cells per step^2 at a centimetre a cell and 60
# steps a second (js/elements.js:30).
GRAVITY = np.float32(9.81 / 3600 / 0.01)


def kinds():
    k = np.zeros(256, dtype=np.uint8)
    for t, kind in _KINDS.items():
        k[t] = kind
    return k
