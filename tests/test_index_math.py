"""CPU checks of the integer shortcuts the kernels take in place of integer division.

The 3D G2P stages its LDS tile and the 3D P2G empties its tile by numbering a box's nodes
t = x + ex (y + ey z) and recovering (x, y, z) as floor((t + 1/2) * fl(1/e)) in binary32, with
fl(1/e) the correctly rounded reciprocal. That equals t // e for every t below 4096 and e up to
2048 (the tiles hold at most 2,400 nodes, the P2G tile 1,024). The P2G marks the bricks of a
one-brick pass, at most two an axis, by the bits of the brick's number. Both are checked here
exhaustively over the ranges the kernels use, with numpy's binary32 arithmetic, which rounds as
the GPU's v_mul_f32 and correctly rounded division do.
"""
import numpy as np


def test_float_reciprocal_floor_is_integer_division():
    t = np.arange(4096, dtype=np.int64)
    tf = t.astype(np.float32) + np.float32(0.5)
    for e in range(1, 2049):
        inv = np.float32(1.0) / np.float32(e)
        got = (tf * inv).astype(np.int64)  # truncation of a positive value: the floor
        assert np.array_equal(got, t // e), e


def test_float_reciprocal_recovers_box_coordinates():
    # every box the G2P tile holds (up to 20 x 12 x 10 nodes): t -> (x, y, z) round trip
    for ex in range(1, 21):
        for ey in range(1, 13):
            nodes = ex * ey * 10
            t = np.arange(nodes, dtype=np.int64)
            inv_ex, inv_ey = np.float32(1.0) / np.float32(ex), np.float32(1.0) / np.float32(ey)
            row = ((t.astype(np.float32) + np.float32(0.5)) * inv_ex).astype(np.int64)
            x = t - row * ex
            z = ((row.astype(np.float32) + np.float32(0.5)) * inv_ey).astype(np.int64)
            y = row - z * ey
            assert np.array_equal(x + ex * (y + ey * z), t)
            assert x.min() >= 0 and x.max() < ex and y.min() >= 0 and y.max() < ey


def test_brick_bits_match_division():
    # a box reaching e bricks an axis, e in {1, 2}: brick t's offsets as bits of t (x fastest)
    for ex in (1, 2):
        for ey in (1, 2):
            for ez in (1, 2):
                sx, sy = ex - 1, ey - 1
                for t in range(ex * ey * ez):
                    bits = (t & sx, (t >> sx) & sy, t >> (sx + sy))
                    div = (t % ex, (t // ex) % ey, t // (ex * ey))
                    assert bits == div
