"""Digests of the step fixtures (tests/golden/steps_*.json), shared by the
fixture writer and the GPU tests that check against them."""
import hashlib

import numpy as np


def cells_digest(cells) -> str:
    """SHA-256 of a uint8 grid (C order [z, y, x]) and its shape."""
    a = np.ascontiguousarray(cells, dtype=np.uint8)
    h = hashlib.sha256(np.asarray(a.shape, "<i8").tobytes())
    h.update(a.tobytes())
    return h.hexdigest()


def fields_digest(fields) -> str:
    """SHA-256 over every orientation's int32 legal overlap field (C order
    [lz, ly, lx], little-endian, each with its shape; an orientation whose
    kernel does not fit contributes a -1 marker)."""
    h = hashlib.sha256()
    for f in fields:
        if f is None:
            h.update(np.asarray([-1], "<i8").tobytes())
            continue
        a = np.ascontiguousarray(f, dtype="<i4")
        h.update(np.asarray(a.shape, "<i8").tobytes())
        h.update(a.tobytes())
    return h.hexdigest()
