"""Checkpoint files (dgr.checkpoint) against the reference's rules in
BSSN_GR/src/bssnCtx.cpp:783-1000: the alternating index (0 when step is a
multiple of 2 freq, else 1; 3 once at the merger), the .cp keys and values,
restore from the index with the larger step (0 on a tie), and a bit-exact
round trip of the block table and the padded state."""
import json

import numpy as np
import pytest

from dgr import checkpoint as CK
from dgr import workloads as W


def test_index_alternates_and_marks_the_merger():
    far = ((3.0, 0.0, 0.0), (-3.0, 0.0, 0.0))
    near = ((0.02, 0.0, 0.0), (-0.03, 0.0, 0.0))
    assert CK.checkpoint_index(0, 50, far, False) == (0, False)
    assert CK.checkpoint_index(50, 50, far, False) == (1, False)
    assert CK.checkpoint_index(100, 50, far, False) == (0, False)
    assert CK.checkpoint_index(150, 50, near, False) == (3, True)   # first checkpoint after the merger
    assert CK.checkpoint_index(200, 50, near, True) == (0, True)    # later ones alternate again


def _write(prefix, idx, step, state, t):
    CK.write_checkpoint(prefix, idx, state, t, time_begin=0.0, time_end=100.0, time=0.25 * step, step=step,
                        dt=0.25, bh=((3.0, 0.1, 0.0), (-3.0, -0.1, 0.0)))


def test_round_trip_and_cp_keys(tmp_path):
    t = W.puncture_octree(0.002)
    st = np.random.default_rng(4).standard_normal((24, W.total_padded(t)))
    prefix = str(tmp_path / "bssn_cp")
    _write(prefix, 1, 75, st, t)
    with open(prefix + "_step_1.cp") as f:
        cp = json.load(f)
    assert tuple(cp) == CK.CP_KEYS
    assert cp["DENDRO_TS_STEP_CURRENT"] == 75 and cp["DENDRO_TS_TIME_CURRENT"] == 18.75
    assert cp["DENDRO_TS_LAST_IO_TIME"] == cp["DENDRO_TS_TIME_CURRENT"]
    assert cp["DENDRO_TS_NUM_VARS"] == 24 and cp["DENDRO_TS_ELEMENT_ORDER"] == 6
    assert (cp["DENDRO_BH1_X"], cp["DENDRO_BH2_Y"]) == (3.0, -0.1)
    idx, cp2, blocks, state = CK.restore_checkpoint(prefix)
    assert idx == 1 and cp2 == cp
    assert blocks.tobytes() == np.ascontiguousarray(t).tobytes()
    assert np.array_equal(state, st)


def test_restore_takes_the_larger_step(tmp_path):
    t = W.single_block_19()
    a = np.zeros((24, W.total_padded(t)))
    b = np.ones((24, W.total_padded(t)))
    prefix = str(tmp_path / "run")
    _write(prefix, 0, 100, a, t)
    _write(prefix, 1, 150, b, t)
    assert CK.restore_checkpoint(prefix)[0] == 1 and CK.restore_checkpoint(prefix)[3][0, 0] == 1.0
    _write(prefix, 0, 200, a, t)
    assert CK.restore_checkpoint(prefix)[0] == 0
    _write(prefix, 1, 200, b, t)           # a tie goes to index 0
    assert CK.restore_checkpoint(prefix)[0] == 0


def test_missing_or_foreign_files(tmp_path):
    prefix = str(tmp_path / "none")
    with pytest.raises(FileNotFoundError):
        CK.restore_checkpoint(prefix)
    t = W.single_block_19()
    _write(prefix, 0, 10, np.zeros((24, W.total_padded(t))), t)
    with open(prefix + "_0_0.var", "r+b") as f:
        f.write(b"NOTAVAR!")
    with pytest.raises(ValueError):
        CK.restore_checkpoint(prefix)
