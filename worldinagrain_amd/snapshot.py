"""World snapshots in the reference's container (js/snapshot.js:1-19,97-192,512-545).

A snapshot is gzipped: the magic "SANDSNAP", the header's length in bytes
(u32, little-endian), the header (JSON), then each part's bytes in the
header's order, each padded to 4 with zeros.  The header holds the format
(1), the world's size, the element ids, the simulation's state and each
part's name and size (saveSnapshot, js/snapshot.js:97-160); the JSON is
written as JSON.stringify writes it: no spaces, keys in order, numbers as
JavaScript prints them.

The parts are the liquid path's buffers under the reference's names
(js/sim.js:1429-1438, a pair as name0 and name1): the current cells buffer
as "cells<cur>", "particles0", "particles1", "counters", "field", "blocks",
"fluidCg", and this library's "solve" (its FP64 solve state, which the
reference does not have; a snapshot without it loads it empty).  A
reference snapshot also carries the other subsystems' parts (air, heat,
bodies ...): they lie outside this path and are passed over on load.

take()/put() move a world's state as (header, parts) without the
container, which is what serving's eviction keeps; save()/load() add it.
"""
import gzip
import json
import math
import struct
from decimal import Decimal

import numpy as np

from . import abi

MAGIC = b"SANDSNAP"
FORMAT = 1
# The element ids of the liquid step's table (js/elements.js:416-457, the
# table's order: AIR 0 .. GRAVEL 5).
ELEMENT_IDS = ["AIR", "WALL", "SAND", "WATER", "STONE", "GRAVEL"]
# Parts this library keeps that the reference's world does not: loaded
# empty when missing (as js/snapshot.js:31-43's ADDED parts do).
ADDED = {"solve"}


class SnapshotError(ValueError):
    pass


def _align4(n):
    return (n + 3) // 4 * 4


# ------------------------------------------------------------ JSON as JS writes it
def js_number(x):
    """Number::toString(10) of ECMA-262 (the shortest digits that read back
    as x, placed by the exponent), as JSON.stringify writes a number; NaN
    and the infinities are null."""
    if isinstance(x, bool):
        return "true" if x else "false"
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    x = float(x)
    if math.isnan(x) or math.isinf(x):
        return "null"
    if x == 0.0:
        return "0"
    if x < 0.0:
        return "-" + js_number(-x)
    _, digits, exp = Decimal(repr(x)).as_tuple()     # repr: the shortest digits that read back as x
    ds = "".join(map(str, digits))
    s = ds.rstrip("0")
    exp += len(ds) - len(s)                            # trailing zeros into the exponent
    ds = s.lstrip("0")
    k = len(ds)
    n = exp + k        # x = 0.d1...dk x 10^n
    if k <= n <= 21:
        return ds + "0" * (n - k)
    if 0 < n <= 21:
        return ds[:n] + "." + ds[n:]
    if -6 < n <= 0:
        return "0." + "0" * (-n) + ds
    e = n - 1
    mant = ds if k == 1 else ds[0] + "." + ds[1:]
    return mant + "e" + ("+" if e >= 0 else "-") + str(abs(e))


def js_json(v):
    """JSON.stringify(v): no whitespace, keys in insertion order."""
    if v is None:
        return "null"
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, float, np.integer, np.floating)):
        return js_number(v)
    if isinstance(v, str):
        return json.dumps(v, ensure_ascii=False)
    if isinstance(v, dict):
        return "{" + ",".join(json.dumps(str(k), ensure_ascii=False) + ":" + js_json(x) for k, x in v.items()) + "}"
    if isinstance(v, (list, tuple)):
        return "[" + ",".join(js_json(x) for x in v) + "]"
    raise TypeError(f"not JSON: {type(v).__name__}")


# ------------------------------------------------------------ the container
def pack_payload(header, parts):
    """The bytes saveSnapshot gzips (js/snapshot.js:139-157), as a
    memoryview: header["parts"] gives the order and sizes; parts maps each
    name to its bytes."""
    head = js_json(header).encode("utf-8")
    views = []
    for p in header["parts"]:
        data = memoryview(parts[p["name"]]).cast("B")
        if data.nbytes != p["bytes"]:
            raise SnapshotError(f"part {p['name']}: {data.nbytes} bytes, the header says {p['bytes']}")
        views.append(data)
    start = len(MAGIC) + 4 + _align4(len(head))
    out = np.empty(start + sum(_align4(v.nbytes) for v in views), dtype=np.uint8)   # (no fill: the pads are zeroed below)
    out[:start] = 0
    out[:len(MAGIC)] = np.frombuffer(MAGIC, dtype=np.uint8)
    out[len(MAGIC):len(MAGIC) + 4] = np.frombuffer(struct.pack("<I", len(head)), dtype=np.uint8)
    out[len(MAGIC) + 4:len(MAGIC) + 4 + len(head)] = np.frombuffer(head, dtype=np.uint8)
    at = start
    for v in views:
        out[at:at + v.nbytes] = np.frombuffer(v, dtype=np.uint8)
        out[at + v.nbytes:at + _align4(v.nbytes)] = 0
        at += _align4(v.nbytes)
    return out.data


def pack(header, parts, level=6):
    return gzip.compress(pack_payload(header, parts), compresslevel=level, mtime=0)


def unpack_payload(raw):
    """readSnapshot after its gunzip (js/snapshot.js:170-192): (header,
    {name: bytes}); the same refusals, in the same words."""
    raw = memoryview(raw).cast("B")
    if raw.nbytes < len(MAGIC) + 4 or bytes(raw[:len(MAGIC)]) != MAGIC:
        raise SnapshotError("not a snapshot")
    (length,) = struct.unpack_from("<I", raw, len(MAGIC))
    header = json.loads(bytes(raw[len(MAGIC) + 4:len(MAGIC) + 4 + length]).decode("utf-8"))
    if header.get("format") != FORMAT:
        raise SnapshotError(f"saved in snapshot format {header.get('format')}; this version reads format {FORMAT}")
    parts = {}
    at = len(MAGIC) + 4 + _align4(length)
    for p in header["parts"]:
        if at + p["bytes"] > raw.nbytes:
            raise SnapshotError("the snapshot is cut short")
        parts[p["name"]] = raw[at:at + p["bytes"]]
        at += _align4(p["bytes"])
    return header, parts


def unpack(blob):
    try:
        raw = gzip.decompress(blob)
    except (OSError, EOFError) as e:
        raise SnapshotError("not a snapshot") from e
    return unpack_payload(raw)


# ------------------------------------------------------------ a world's state
def _part_name(p, cur=0):
    return f"cells{cur}" if p == abi.PART_CELLS else abi.PART_NAMES[p]


def take(sim, elements=None, extra=None):
    """The world's (header, parts) as saveSnapshot gathers them (js/sim.js:
    1429-1465): its parts by name, and the state the liquid keeps under
    snapshotState's names (plus dt, the liquid's step, which the reference
    takes from its clock).  extra: more header fields, kept as given (a
    reference snapshot's water, gas and bodies)."""
    w = sim.world
    parts, items = {}, []
    for p in range(abi.LIQUID_PARTS):
        name = _part_name(p)
        data = np.ascontiguousarray(w.download(p))
        parts[name] = data
        items.append({"name": name, "bytes": int(data.nbytes)})
    header = {"format": FORMAT, "width": sim.width, "height": sim.height,
              "elements": list(elements or ELEMENT_IDS)}
    header.update(extra or {})
    header["state"] = {
        "cur": 0, "frameNo": int(sim.frame_no), "shiftedFirst": bool(sim.shifted_first), "rng": int(sim.rng_state),
        "metresPerCell": float(sim.dx), "earthGravity": float(sim.earth_g),
        "fluidStats": {"active": int(sim.fluid_stats["active"]), "viscous": int(sim.fluid_stats["viscous"])},
        "fluidCycles": int(sim.fluid_cycles), "dt": float(sim.dt),
    }
    header["parts"] = items
    return header, parts


def put(sim, header, parts, elements=None):
    """restoreSnapshot's checks and loads for the liquid's parts
    (js/snapshot.js:512-545): a world of the snapshot's size, the same
    elements, each part the size this world's is (a part of ADDED may be
    missing: it loads empty); then the state.  Nothing is changed if any
    check fails."""
    w = sim.world
    if header["width"] != sim.width or header["height"] != sim.height:
        raise SnapshotError(f"the snapshot is of a {header['width']}×{header['height']} world, "
                            f"this one is {sim.width}×{sim.height}")
    if ",".join(header["elements"]) != ",".join(elements or ELEMENT_IDS):
        raise SnapshotError("saved with a different set of elements")
    cur = int(header["state"].get("cur", 0))
    load, misfits = [], []
    for p in range(abi.LIQUID_PARTS):
        name, size = _part_name(p, cur), int(w.layout.part_bytes[p])
        data = parts.get(name)
        if data is None and _part_name(p) in ADDED:
            load.append((p, np.zeros(size, dtype=np.uint8)))
        elif data is None or memoryview(data).nbytes != size:
            misfits.append(name)
        else:
            load.append((p, np.frombuffer(memoryview(data).cast("B"), dtype=np.uint8)))
    if misfits:
        raise SnapshotError(f"saved by a version that lays out {', '.join(misfits)} differently")
    for p, data in load:
        w.upload(p, data)
    s = header["state"]
    sim.frame_no = int(s["frameNo"])
    sim.shifted_first = bool(s["shiftedFirst"])
    sim.rng_state = int(s["rng"]) & 0xFFFFFFFF
    sim.dx = float(s["metresPerCell"])
    sim.earth_g = float(s["earthGravity"])
    fs = s.get("fluidStats", {})
    sim.fluid_stats = {"active": int(fs.get("active", 1)), "viscous": int(fs.get("viscous", 0))}
    sim.fluid_cycles = int(s.get("fluidCycles", sim.solver.max_cycles))
    if "dt" in s:
        sim.dt = float(s["dt"])


def save(sim, level=6, elements=None, extra=None):
    """The world as a snapshot file's bytes (saveSnapshot)."""
    return pack(*take(sim, elements, extra), level=level)


def load(sim, blob, elements=None):
    """A snapshot's bytes into sim's world and state (readSnapshot, then
    restoreSnapshot); returns the header."""
    header, parts = unpack(blob)
    put(sim, header, parts, elements)
    return header
