"""Bit-exact checkpoint / restart of a solver context.

Replaces the reference's restart files (mulitphase Solver/doublePrecision/
restart.cu:113-152 writeFieldRestart, 424-497 readField), which store
physicalTime and consVec as %.10e text, so a restart there is not bit-exact.
Here the file holds the raw FP64 state and the loop clock.

File layout (little endian):
  magic   8 bytes  b"FKCKPT01"
  header  int64 n_cells, int64 neq, int64 step, float64 dt, float64 time,
          uint64 checksum of the state words
  state   float64 [n_cells][neq]  (reference consVec order)
The checksum is the wrapping sum of the state as uint64 words, multiplied by
their 1-based position (a dropped, swapped or flipped word changes it).
"""
import struct

import numpy as np

MAGIC = b"FKCKPT01"
_HEADER = struct.Struct("<qqqddQ")


class CheckpointError(RuntimeError):
    pass


def checksum(U):
    """Position-weighted wrapping sum of the state's 64-bit words."""
    w = np.ascontiguousarray(U, dtype=np.float64).view(np.uint64).ravel()
    with np.errstate(over="ignore"):
        pos = np.arange(1, w.size + 1, dtype=np.uint64)
        return int(np.sum(w * pos, dtype=np.uint64))


def save(path, solver):
    """Write the solver's state and clock to `path`."""
    U = solver.download_state()
    dt, time, step = solver.clock()
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(_HEADER.pack(U.shape[0], U.shape[1], step, dt, time, checksum(U)))
        U.astype("<f8", copy=False).tofile(f)


def read(path):
    """(U, dt, time, step) of a checkpoint file, checksum verified."""
    with open(path, "rb") as f:
        if f.read(len(MAGIC)) != MAGIC:
            raise CheckpointError(f"{path}: not a checkpoint file")
        hdr = f.read(_HEADER.size)
        if len(hdr) != _HEADER.size:
            raise CheckpointError(f"{path}: truncated header")
        nc, neq, step, dt, time, ck = _HEADER.unpack(hdr)
        U = np.fromfile(f, dtype="<f8", count=nc * neq)
    if U.size != nc * neq:
        raise CheckpointError(f"{path}: truncated state ({U.size} of {nc * neq} values)")
    U = U.reshape(nc, neq)
    if checksum(U) != ck:
        raise CheckpointError(f"{path}: checksum mismatch")
    return U, dt, time, step


def load(path, solver):
    """Restore state and clock into `solver` (same mesh and model)."""
    U, dt, time, step = read(path)
    if U.shape != (solver.n_cells, solver.neq):
        raise CheckpointError(f"{path}: state {U.shape} does not match the solver "
                              f"({solver.n_cells}, {solver.neq})")
    solver.upload_state(U)
    solver.set_clock(dt, time, step)
    return dt, time, step
