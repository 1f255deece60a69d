"""Host producers of the per-frame inputs: the camera block
(camera.rs:33-183, render/state/mod.rs:574-616) and the terrain clipmap
(clipmap.rs:62-122).  Checked against geometry the construction fixes."""
from __future__ import annotations

import math

import numpy as np


def _rays(cam, x, y):
    return (np.array(cam.ray00[:3], np.float64) + x * np.array(cam.ray_dx[:3], np.float64)
            + y * np.array(cam.ray_dy[:3], np.float64))


def test_centre_ray_is_the_view_direction(vx):
    for yaw, pitch in ((0.0, 0.0), (1.1, -0.3), (-2.5, 0.7)):
        pose = vx.make_pose((3.0, 10.0, -4.0), yaw, pitch, 70.0)
        w, h = 640, 360
        cam = vx.camera_build(pose, (w, h), (w, h), 5)
        fwd = np.array([math.sin(yaw) * math.cos(pitch), math.sin(pitch), math.cos(yaw) * math.cos(pitch)])
        # pixel centre (x + 0.5) with x = w/2 - 0.5 is the image centre
        d = _rays(cam, w / 2 - 0.5, h / 2 - 0.5)
        d /= np.linalg.norm(d)
        assert np.allclose(d, fwd, atol=1e-5)
        # vertical field of view: top and bottom edge rays are 70 degrees apart
        top = _rays(cam, w / 2 - 0.5, -0.5)
        bot = _rays(cam, w / 2 - 0.5, h - 0.5)
        ang = math.acos(np.dot(top, bot) / np.linalg.norm(top) / np.linalg.norm(bot))
        assert abs(math.degrees(ang) - 70.0) < 1e-3
        # view_proj maps a point along the centre ray to NDC (0, 0)
        vp = np.array(cam.view_proj[:], np.float64).reshape(4, 4).T  # column-major storage
        p = np.array([3.0, 10.0, -4.0]) + 25.0 * fwd
        clip = vp @ np.append(p, 1.0)
        assert np.allclose(clip[:2] / clip[3], 0.0, atol=1e-5) and 0.0 < clip[2] / clip[3] < 1.0
        assert cam.max_steps == 128 and cam.chunk_size == 64 and cam.voxel_params[0] == np.float32(0.1)


def test_jitter_is_halton_2_3(vx):
    pose = vx.make_pose((0.0, 5.0, 0.0), 0.4, 0.1, 70.0)
    w, h = 320, 180
    base = vx.camera_build(pose, (w, h), (w, h), 0)
    r00 = np.array(base.ray00[:3], np.float64)
    dx, dy = np.array(base.ray_dx[:3], np.float64), np.array(base.ray_dy[:3], np.float64)

    def halton(i, b):
        f, r = 1.0, 0.0
        while i > 0:
            f /= b
            r += f * (i % b)
            i //= b
        return r

    for frame in range(6):
        cam = vx.camera_build(pose, (w, h), (w, h), frame, jitter=True)
        jx, jy = halton(frame + 1, 2) - 0.5, halton(frame + 1, 3) - 0.5
        # the jittered first ray equals the unjittered plane sampled at (jx, jy);
        # the ray plane is affine in pixel coordinates at the far plane
        want = r00 + jx * dx + jy * dy
        assert np.allclose(np.array(cam.ray00[:3]), want, atol=2e-5)


def test_clipmap_toroidal_storage(vx):
    p = vx.world_params(42)
    cx, cz = 123.4, -57.9
    clip, heights = vx.clipmap_build(p, cx, cz, levels=3, res=64, base_cell_m=1.0)
    assert clip.levels == 3 and clip.res == 64
    for lvl in range(3):
        ox, oz, cell = clip.level[lvl][0], clip.level[lvl][1], clip.level[lvl][2]
        assert cell == 2.0 ** lvl
        # snapped origin: camera sits res/2 cells from the origin, on a cell boundary
        assert ox == math.floor((cx - 32 * cell) / cell) * cell and oz == math.floor((cz - 32 * cell) / cell) * cell
        offx, offz = clip.offset[lvl][0], clip.offset[lvl][1]
        assert offx == int(ox / cell) % 64 and offz == int(oz / cell) % 64
        for tz in (0, 7, 31, 63):
            for tx in (0, 13, 40, 63):
                wx = np.float32(ox) + (np.float32(tx) + np.float32(0.5)) * np.float32(cell)
                wz = np.float32(oz) + (np.float32(tz) + np.float32(0.5)) * np.float32(cell)
                g = vx.ground_height(p, int(math.floor(wx / np.float32(0.1))), int(math.floor(wz / np.float32(0.1))))
                got = heights[lvl, (tz + offz) % 64, (tx + offx) % 64]
                assert got == np.float32(g) * np.float32(0.1)
    # moving the camera by one cell shifts the origin by one cell and keeps
    # every shared texel at the same storage address
    clip2, heights2 = vx.clipmap_build(p, cx + 1.0, cz, levels=1, res=64)
    assert clip2.level[0][0] == clip.level[0][0] + 1.0
    assert clip2.offset[0][0] == (clip.offset[0][0] + 1) % 64
    moved = (clip.offset[0][0] + 0) % 64  # logical column 0 of the old map is gone
    keep = np.ones(64, bool)
    keep[moved] = False
    assert np.array_equal(heights2[0][:, keep], heights[0][:, keep])
