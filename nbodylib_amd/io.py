"""The reference driver's raw particle file, parts.out (src/tests/test_kdtree.cxx:60-73 writes it, :181-208 reads
it): one record of six native-endian doubles per particle, x y z vx vy vz (Particle::GetPhase(0..5)), no header.
Reading sets every mass to 1 and the IDs to the input order, as the driver does."""
import os

import numpy as np

RECORD = 6 * 8  # bytes per particle


def read_parts_out(path, n=None):
    """(pos (n, 3), vel (n, 3), mass (n,) of ones).  n: read the first n records (the driver's size argument);
    None reads them all."""
    size = os.path.getsize(path)
    if size % RECORD:
        raise ValueError(f"{path}: {size} bytes is not a whole number of 48-byte records")
    total = size // RECORD
    n = total if n is None else int(n)
    if n > total:
        raise ValueError(f"{path} holds {total} particles, {n} asked for")
    ph = np.fromfile(path, dtype=np.float64, count=6 * n).reshape(n, 6)
    return np.ascontiguousarray(ph[:, :3]), np.ascontiguousarray(ph[:, 3:]), np.ones(n)


def write_parts_out(path, pos, vel=None):
    """Write pos (n, 3) and vel (n, 3; zeros when None) as parts.out records."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    vel = np.zeros_like(pos) if vel is None else np.asarray(vel, dtype=np.float64).reshape(-1, 3)
    if vel.shape != pos.shape:
        raise ValueError("pos and vel must both be (n, 3)")
    np.concatenate([pos, vel], axis=1).astype(np.float64).tofile(path)
