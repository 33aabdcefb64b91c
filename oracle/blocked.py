"""Placed-copy blocked field of the CPU oracle -- test infrastructure only
(imported by tests/ as the checker; the product path never uses it).

Restates BlockedField (engine/pack_geometry/include/spectrapack/geometry/
conservative_fields.hpp:127-164, engine/pack_geometry/src/conservative_fields.cpp:
1666-1887): an immutable mask, a uint32 placed count per cell, and each
copy's footprint by copy id.  add() counts a footprint in (FIELD_COPY_ID for
an empty or present id, FIELD_COUNT_OVERFLOW at 2^32 - 1), remove() counts it
out (FIELD_COPY_ID for an absent id, FIELD_COUNT_UNDERFLOW when a count is
already 0), blocked() is mask or count != 0 (conservative_fields.cpp:1666-1679).
"""
import numpy as np


class FieldFailure(Exception):
    def __init__(self, code):
        super().__init__(code)
        self.code = code


class BlockedField:
    def __init__(self, mask):
        self.mask = np.ascontiguousarray(mask, dtype=np.uint8).copy()
        self.counts = np.zeros(self.mask.shape, np.uint32)
        self.footprints = {}

    def add(self, copy_id, footprint):
        """conservative_fields.cpp:1698-1797: footprint is a {0,1} grid of the window's shape."""
        fp = np.asarray(footprint) != 0
        if fp.shape != self.mask.shape:
            raise FieldFailure("FIELD_GRID")
        if copy_id == "" or copy_id is None:
            raise FieldFailure("FIELD_COPY_ID")
        if copy_id in self.footprints:
            raise FieldFailure("FIELD_COPY_ID")
        if (self.counts[fp] == np.iinfo(np.uint32).max).any():
            raise FieldFailure("FIELD_COUNT_OVERFLOW")
        idx = np.flatnonzero(fp)
        self.counts.flat[idx] += 1
        self.footprints[copy_id] = idx

    def remove(self, copy_id):
        """conservative_fields.cpp:1805-1887."""
        if copy_id not in self.footprints:
            raise FieldFailure("FIELD_COPY_ID")
        idx = self.footprints[copy_id]
        if (self.counts.flat[idx] == 0).any():
            raise FieldFailure("FIELD_COUNT_UNDERFLOW")
        self.counts.flat[idx] -= 1
        del self.footprints[copy_id]

    def placed_count(self, zyx):
        return int(self.counts[tuple(zyx)])

    def blocked(self):
        return ((self.mask != 0) | (self.counts != 0)).astype(np.uint8)
