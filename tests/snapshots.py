"""Read/write the drivers' binary block snapshots (int32 count, then per
block int32 nx, ny and float32 h, hu, hv, b in Float2D order)."""
from __future__ import annotations

import numpy as np


def write(path, blocks) -> None:
    with open(path, "wb") as f:
        np.array([len(blocks)], np.int32).tofile(f)
        for bl in blocks:
            np.array([bl["nx"], bl["ny"]], np.int32).tofile(f)
            for k in ("h", "hu", "hv", "b"):
                np.ascontiguousarray(bl[k], np.float32).tofile(f)


def read(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    off = 4
    out = []
    for _ in range(n):
        nx, ny = (int(v) for v in np.frombuffer(raw, np.int32, 2, off))
        off += 8
        m = (nx + 2) * (ny + 2)
        arrs = []
        for _k in range(4):
            arrs.append(np.frombuffer(raw, np.float32, m, off).reshape(nx + 2, ny + 2).copy())
            off += 4 * m
        out.append(dict(nx=nx, ny=ny, h=arrs[0], hu=arrs[1], hv=arrs[2], b=arrs[3]))
    return out
