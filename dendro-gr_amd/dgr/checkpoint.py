"""Checkpoints of a run on the padded blocks, in the file set and with the
alternation of the reference's BSSNCtx::write_checkpt / restore_checkpt
(BSSN_GR/src/bssnCtx.cpp:783-862 and 864-1000):

  <prefix>_step_<i>.cp        JSON, the reference's keys (DENDRO_TS_*, the BH
                              positions), written by rank 0 only
  <prefix>_<i>_<rank>.var     the 24 evolved variables of this rank
  <prefix>_octree_<i>_<rank>.oct   this rank's block table

The index i alternates between 0 and 1 (step % (2 freq) == 0 -> 0, else 1,
bssnCtx.cpp:786-789), and is 3 for the one checkpoint written when the two
black holes first come within 0.1 of each other (:791-796). Restoring reads
the .cp of both indices 0 and 1 and takes the one with the larger step, index
0 on a tie (:892-946).

The .var and .oct payloads are written by dendrolib in the reference
(io::checkpoint::writeVecToFile / writeOctToFile, not in this checkout): there
they hold the zipped nodal vector and the octants. Here the state lives in the
padded blocks, so the payloads are this build's own, little-endian:
  .var  b"DGRVAR01", uint32 number of variables, uint64 szpdof, then each
        variable's szpdof doubles (padding included, so a restored state
        needs no padding refresh before its first right-hand side);
  .oct  b"DGROCT01", uint64 number of blocks, then the dgr_block records
        (include/dgr_bssn.h: offset, sz[3], bflag, pmin[3], pmax[3]).
"""
from __future__ import annotations

import json
import os

import numpy as np

from .workloads import BLOCK_DTYPE

# the keys of the .cp file, in the order bssnCtx.cpp:829-854 writes them
CP_KEYS = (
    "DENDRO_TS_TIME_BEGIN", "DENDRO_TS_TIME_END", "DENDRO_TS_ELEMENT_ORDER", "DENDRO_TS_TIME_CURRENT",
    "DENDRO_TS_STEP_CURRENT", "DENDRO_TS_TIME_STEP_SIZE", "DENDRO_TS_LAST_IO_TIME",
    "DENDRO_TS_WAVELET_TOLERANCE", "DENDRO_TS_LOAD_IMB_TOLERANCE", "DENDRO_TS_NUM_VARS",
    "DENDRO_TS_ACTIVE_COMM_SZ", "DENDRO_BH1_X", "DENDRO_BH1_Y", "DENDRO_BH1_Z", "DENDRO_BH2_X",
    "DENDRO_BH2_Y", "DENDRO_BH2_Z",
)
VAR_MAGIC = b"DGRVAR01"
OCT_MAGIC = b"DGROCT01"


def checkpoint_index(step: int, freq: int, bh, merged_written: bool) -> tuple[int, bool]:
    """(index, merged_written after this checkpoint) as bssnCtx.cpp:786-796."""
    idx = 0 if step % (2 * freq) == 0 else 1
    merged = float(np.linalg.norm(np.subtract(bh[0], bh[1]))) < 0.1
    if merged and not merged_written:
        return 3, True
    return idx, merged_written


def _paths(prefix: str, idx: int, rank: int):
    return (f"{prefix}_step_{idx}.cp", f"{prefix}_{idx}_{rank}.var", f"{prefix}_octree_{idx}_{rank}.oct")


def write_checkpoint(prefix: str, idx: int, state, blocks: np.ndarray, *, time_begin: float, time_end: float,
                     time: float, step: int, dt: float, bh, element_order: int = 6, wavelet_tol: float = 1e-4,
                     load_imb_tol: float = 0.1, comm_size: int = 1, rank: int = 0) -> None:
    """Writes checkpoint `idx` (see checkpoint_index) of this rank: its block
    table and its 24 x szpdof state (a numpy array or a device tensor, copied
    to the host), and, on rank 0, the .cp file."""
    cp, var, oct_ = _paths(prefix, idx, rank)
    host = state.detach().cpu().numpy() if hasattr(state, "detach") else np.asarray(state)
    host = np.ascontiguousarray(host, dtype="<f8")
    if host.ndim != 2:
        raise ValueError("state must be (number of variables, szpdof)")
    blocks = np.ascontiguousarray(blocks, dtype=BLOCK_DTYPE)
    with open(oct_, "wb") as f:
        f.write(OCT_MAGIC)
        f.write(np.uint64(len(blocks)).tobytes())
        f.write(blocks.tobytes())
    with open(var, "wb") as f:
        f.write(VAR_MAGIC)
        f.write(np.uint32(host.shape[0]).tobytes())
        f.write(np.uint64(host.shape[1]).tobytes())
        f.write(host.tobytes())
    if rank == 0:
        vals = (time_begin, time_end, element_order, time, step, dt, time, wavelet_tol, load_imb_tol,
                host.shape[0], comm_size, *[float(c) for p in bh for c in p])
        with open(cp, "w") as f:
            json.dump(dict(zip(CP_KEYS, vals)), f, indent=4)
            f.write("\n")


def read_cp(prefix: str) -> tuple[int, dict] | None:
    """The .cp to restore from: of indices 0 and 1 the one with the larger
    DENDRO_TS_STEP_CURRENT (0 on a tie); None when neither exists."""
    found = {}
    for idx in (0, 1):
        p = f"{prefix}_step_{idx}.cp"
        if os.path.exists(p):
            with open(p) as f:
                found[idx] = json.load(f)
    if not found:
        return None
    steps = {i: int(found[i]["DENDRO_TS_STEP_CURRENT"]) if i in found else 0 for i in (0, 1)}
    idx = 1 if steps[0] < steps[1] else 0
    if idx not in found:
        return None
    return idx, found[idx]


def restore_checkpoint(prefix: str, rank: int = 0):
    """(index, .cp contents, block table, state as a (vars, szpdof) numpy
    array) of the checkpoint read_cp selects; FileNotFoundError if there is
    none, ValueError if a payload is not in this format."""
    sel = read_cp(prefix)
    if sel is None:
        raise FileNotFoundError(f"no checkpoint {prefix}_step_{{0,1}}.cp")
    idx, cp = sel
    _, var, oct_ = _paths(prefix, idx, rank)
    with open(oct_, "rb") as f:
        if f.read(8) != OCT_MAGIC:
            raise ValueError(f"{oct_}: not a block-table checkpoint")
        nb = int(np.frombuffer(f.read(8), "<u8")[0])
        blocks = np.frombuffer(f.read(nb * BLOCK_DTYPE.itemsize), BLOCK_DTYPE).copy()
    with open(var, "rb") as f:
        if f.read(8) != VAR_MAGIC:
            raise ValueError(f"{var}: not a variable checkpoint")
        nv = int(np.frombuffer(f.read(4), "<u4")[0])
        n = int(np.frombuffer(f.read(8), "<u8")[0])
        state = np.frombuffer(f.read(nv * n * 8), "<f8").reshape(nv, n).copy()
    if nv != int(cp["DENDRO_TS_NUM_VARS"]):
        raise ValueError("the .var holds another number of variables than the .cp names")
    return idx, cp, blocks, state
