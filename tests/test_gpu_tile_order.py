"""Tile order of multi-sample frames: once a frame set has rendered a frame,
its next frame hands out 8x8 tiles heaviest first by the times the previous
one recorded.  Nine consecutive frames (every one of the four frame sets
used twice or more, so the later frames run in a history-based order) must
accumulate to the oracle's image bit for bit: through the camera-ray queue
(a frame of whole tiles) and through the feature form (a ragged frame with
depth of field)."""
import numpy as np
import pytest

import dmt
import oracle_ctypes as orc

pytestmark = pytest.mark.gpu

FRAMES = 9


def _accumulate(sd, f, frames):
    osc = orc.OracleScene.from_scene_data(sd)
    acc = None
    for k in range(frames):
        g = f.copy()
        g.frame_index = f.frame_index + k
        g.reset_accum = 1 if k == 0 else 0
        acc, _ = osc.render(g, 32, 0, accum=acc)
    return acc


def _check(device, sd, f, label):
    sc = dmt.Scene(device, sd)
    sc.render_frames(f, FRAMES)
    got = sc.readback(f.width, f.height)
    sc.close()
    want = _accumulate(sd, f, FRAMES)
    same = float(np.mean(np.all(got == want, axis=-1)))
    print(f"{label}: equal to the oracle on {same:.6f} of pixels", flush=True)
    assert np.array_equal(got, want), label


def test_queued_frames_in_tile_order_equal_the_oracle(device):
    sd = dmt.generate(0, seed=41, width=64, height=48)
    f = sd.frame.copy()
    f.spp = 3
    _check(device, sd, f, "queued, 64x48, 3 spp, 9 frames")


def test_feature_frames_in_tile_order_equal_the_oracle(device):
    sd = dmt.generate(0, seed=42, width=60, height=44)
    f = sd.frame.copy()
    f.spp = 2
    f.dof[:] = (0.05, 10.0, 0.0, 0.0)
    _check(device, sd, f, "features, 60x44, 2 spp, 9 frames")
