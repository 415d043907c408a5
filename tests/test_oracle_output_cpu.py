"""CPU pins of the numpy output-encoding oracle: the u8 rule and the orientation index
maps checked on hand-worked cases (extern_pipeline.rs:381-383, output.rs:31-70)."""
import numpy as np

import oracle


def test_u8_rule_hand_cases():
    x = np.array([[[-1.0, 0.0, 0.5]], [[1.0, 2.0, np.nan]], [[1 / 255, 254.99 / 255, 0.999]]], np.float32)
    assert oracle.encode_u8(x).reshape(-1).tolist() == [0, 0, 127, 255, 255, 0, 1, 254, 254]


def test_orientation_maps():
    # 2 wide, 3 tall; pixel value = 10 y + x
    a = (10 * np.arange(3)[:, None] + np.arange(2)[None, :]).astype(np.float32)[..., None].repeat(3, 2)
    r90 = oracle.orient(a, 1)[..., 0]      # clockwise: the bottom-left pixel lands top-left
    assert r90.tolist() == [[20, 10, 0], [21, 11, 1]]
    assert oracle.orient(a, 2)[..., 0].tolist() == [[21, 20], [11, 10], [1, 0]]
    assert oracle.orient(a, 3)[..., 0].tolist() == [[1, 11, 21], [0, 10, 20]]
    for o in (1, 2, 3):  # four quarter turns are the identity; 90 then 270 too
        assert np.array_equal(oracle.orient(oracle.orient(a, o), (4 - o) % 4), a)
