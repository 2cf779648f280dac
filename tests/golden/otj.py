"""Reader of the reference's trajectory format (docs/formats/trajectory.md).

Header: magic ``ORRERYTJ``, u32 version, u32 flags, u64 N, R timestep, N masses,
u64 FNV-1a checksum; frames: u64 step, R time, x[N], y[N], z[N] (+ v if flag
bit 1), u64 checksum. Little-endian; R = 8 unless flag bit 0.
"""
import struct

import numpy as np

FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3


def fnv1a(data: bytes) -> int:
    h = FNV_OFFSET
    for b in data:
        h ^= b
        h = (h * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def read(path):
    buf = open(path, "rb").read()
    if buf[:8] != b"ORRERYTJ":
        raise ValueError("not a trajectory")
    version, flags, n = struct.unpack_from("<IIQ", buf, 8)
    R = 4 if flags & 1 else 8
    dt_ = "<f4" if R == 4 else "<f8"
    timestep = float(np.frombuffer(buf, dtype=dt_, count=1, offset=24)[0])
    masses = np.frombuffer(buf, dtype=dt_, count=n, offset=24 + R).astype(np.float64)
    off = 24 + R + n * R
    (csum,) = struct.unpack_from("<Q", buf, off)
    if csum != fnv1a(buf[:off]):
        raise ValueError("header checksum mismatch")
    off += 8
    frames = []
    with_v = bool(flags & 2)
    while off < len(buf):
        start = off
        (step,) = struct.unpack_from("<Q", buf, off)
        time = float(np.frombuffer(buf, dtype=dt_, count=1, offset=off + 8)[0])
        off += 8 + R
        arrays = []
        for _ in range(6 if with_v else 3):
            arrays.append(np.frombuffer(buf, dtype=dt_, count=n, offset=off).astype(np.float64))
            off += n * R
        (csum,) = struct.unpack_from("<Q", buf, off)
        if csum != fnv1a(buf[start:off]):
            raise ValueError(f"frame checksum mismatch at step {step}")
        off += 8
        frames.append(dict(step=step, time=time, positions=arrays[:3],
                           velocities=arrays[3:] if with_v else None))
    return dict(version=version, flags=flags, n=n, timestep=timestep, masses=masses, frames=frames)


def read_csv(path):
    import csv
    with open(path) as f:
        rows = list(csv.DictReader(f))
    return [{k: float(v) for k, v in r.items()} for r in rows]
