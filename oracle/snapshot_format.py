"""ORACLE — test infrastructure only (tests/).  A restatement of the
reference's snapshot container, the checker of worldinagrain_amd.snapshot.

write(): saveSnapshot's packing (js/snapshot.js:139-157): MAGIC, the header's
  length as a little-endian u32 (DataView.setUint32(.., true)), the header's
  UTF-8 bytes, zeros to a multiple of 4 (the output starts zeroed), then each
  part at the next offset, its length rounded up to 4.
read(): readSnapshot after the gunzip (js/snapshot.js:170-192), its refusals.

Parity: pinned by the hand-assembled byte strings of tests/test_snapshot.py
(offsets and padding written out there, independently of both)."""
import json
import struct

MAGIC = b"SANDSNAP"


def align(n, to=4):                                   # js/snapshot.js:84
    return -(-n // to) * to


def write(head_bytes, parts):
    """head_bytes: the header's JSON text, encoded; parts: [(bytes)] in the
    header's order."""
    start = len(MAGIC) + 4 + align(len(head_bytes))
    total = start + sum(align(len(p)) for p in parts)
    out = bytearray(total)                            # new Uint8Array(...): zeros
    out[0:len(MAGIC)] = MAGIC
    struct.pack_into("<I", out, len(MAGIC), len(head_bytes))
    out[len(MAGIC) + 4:len(MAGIC) + 4 + len(head_bytes)] = head_bytes
    at = start
    for p in parts:
        out[at:at + len(p)] = p
        at += align(len(p))
    return bytes(out)


def read(raw):
    if len(raw) < len(MAGIC) + 4 or raw[:len(MAGIC)] != MAGIC:
        raise ValueError("not a snapshot")
    (length,) = struct.unpack_from("<I", raw, len(MAGIC))
    header = json.loads(raw[len(MAGIC) + 4:len(MAGIC) + 4 + length].decode("utf-8"))
    if header["format"] != 1:
        raise ValueError(f"saved in snapshot format {header['format']}; this version reads format 1")
    parts = {}
    at = len(MAGIC) + 4 + align(length)
    for part in header["parts"]:
        if at + part["bytes"] > len(raw):
            raise ValueError("the snapshot is cut short")
        parts[part["name"]] = bytes(raw[at:at + part["bytes"]])
        at += align(part["bytes"])
    return header, parts
