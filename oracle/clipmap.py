"""Restatement of the reference's incremental clipmap producer (src/clipmap.rs)
in numpy, binary32 where the reference computes in f32.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

  * Clipmap state (clipmap.rs:44-60): per level the committed origin cell
    (i32::MIN = untouched) and toroidal offset, plus the last update time.
  * update (clipmap.rs:182-577): pass 1 snaps each level's origin to the camera
    and derives first/moved/forced; one cadence gate decides whether the frame
    updates; pass 2 scrolls offsets and emits the uploads (full level, or new
    columns then new rows, each split where it wraps) and the committed params.
  * invalidate_all (clipmap.rs:580-585).
Heights come from a caller-supplied vectorised ground-height function
height_fn(vx, vz) -> voxel y (int arrays), sampled under texel centres
(sample_height_f32, clipmap.rs:93-104).  `wrap_fix` places a strip's spanning
axis at its wrapped storage position instead of from storage 0.

Also here: ground_height_np, a numpy restatement of the seeded synthetic world's
surface height (the library's own generator, binary64 value noise), so that
texel heights can be produced without the product library.
"""
from __future__ import annotations

import numpy as np

F = np.float32
I32_MIN = -(2 ** 31)
EDGE_GUARD = 16  # clipmap.rs:207
VOXEL_M = F(0.1)  # app/config.rs:10


def sat_i32(v) -> np.ndarray:
    """f32 -> i32 like Rust's `as` (saturating, NaN -> 0)."""
    v = np.asarray(v, np.float32)
    out = np.where(np.isnan(v), F(0), v)
    out = np.clip(out.astype(np.float64), -2147483648.0, 2147483647.0)
    return out.astype(np.int64).astype(np.int32)


def level_cell(base_cell: float, level: int) -> np.float32:
    """clipmap.rs:61-64."""
    return F(base_cell) * F(1 << level)


def snapped_origin_cell(cam_x: float, cam_z: float, cell: np.float32, res: int):
    """clipmap.rs:66-79: origin = floor((cam - res/2 * cell) / cell) in cells."""
    half = F(res) * F(0.5)
    ox = F(cam_x) - half * cell
    oz = F(cam_z) - half * cell
    return int(sat_i32(np.floor(ox / cell))), int(sat_i32(np.floor(oz / cell)))


def texel_heights(height_fn, ox_m: np.float32, oz_m: np.float32, cell: np.float32, lxs, lzs) -> np.ndarray:
    """Heights (metres) of logical texels lzs x lxs: ground height of the voxel
    column under each texel centre, times the voxel size (clipmap.rs:93-122)."""
    lxs = np.asarray(lxs, np.int64)
    lzs = np.asarray(lzs, np.int64)
    wx = F(ox_m) + (lxs.astype(np.float32) + F(0.5)) * cell
    wz = F(oz_m) + (lzs.astype(np.float32) + F(0.5)) * cell
    vx = sat_i32(np.floor(wx / VOXEL_M))
    vz = sat_i32(np.floor(wz / VOXEL_M))
    h = np.asarray(height_fn(vx[None, :], vz[:, None]), np.int32)
    return (h.astype(np.float32) * VOXEL_M).astype(np.float32)


class Clipmap:
    """clipmap.rs:44-60 (state) with update (182-577) and invalidate_all (580-585)."""

    def __init__(self, levels: int, res: int, base_cell: float = 1.0, min_interval: float = 0.0,
                 wrap_fix: bool = False):
        self.levels, self.res, self.base_cell = levels, res, base_cell
        self.min_interval, self.wrap_fix = F(min_interval), wrap_fix
        self.invalidate()

    def invalidate(self) -> None:
        self.last_cell = [(I32_MIN, I32_MIN)] * self.levels
        self.tex_offset = [(0, 0)] * self.levels
        self.last_time = F(-np.inf)

    def update(self, height_fn, cam_x: float, cam_z: float, time_s: float):
        res = self.res
        # Pass 1 (clipmap.rs:211-272).
        lv = []
        any_first = any_moved = any_force = False
        for i in range(self.levels):
            cell = level_cell(self.base_cell, i)
            ncx, ncz = snapped_origin_cell(cam_x, cam_z, cell, res)
            ocx, ocz = self.last_cell[i]
            first = ocx == I32_MIN
            dx = 0 if first else ncx - ocx
            dz = 0 if first else ncz - ocz
            moved = first or dx != 0 or dz != 0
            if first:
                force = True
            else:
                cam_lx = int(sat_i32(np.floor((F(cam_x) - F(ocx) * cell) / cell)))
                cam_lz = int(sat_i32(np.floor((F(cam_z) - F(ocz) * cell) / cell)))
                force = (cam_lx < EDGE_GUARD or cam_lz < EDGE_GUARD or cam_lx >= res - EDGE_GUARD
                         or cam_lz >= res - EDGE_GUARD)
            any_first |= first
            any_moved |= moved
            any_force |= force
            lv.append(dict(cell=cell, cx=ncx, cz=ncz, ox=F(ncx) * cell, oz=F(ncz) * cell, first=first, dx=dx, dz=dz,
                           moved=moved))

        # Cadence gate (clipmap.rs:274-284).
        cadence_ok = any_first or (F(time_s) - self.last_time) >= self.min_interval
        do_frame = any_first or any_force or (cadence_ok and any_moved)
        if do_frame:
            self.last_time = F(time_s)

        # Pass 2 (clipmap.rs:289-575).
        uploads = []
        params = dict(levels=self.levels, res=res, base_cell=F(self.base_cell), level=[], offset=[])
        for i, q in enumerate(lv):
            off_x, off_z = self.tex_offset[i]
            cell, ox, oz = q["cell"], q["ox"], q["oz"]

            def up(x, y, lx0, lz0, w, h):
                # Storage rect (x, y, w, h) from logical texel (lx0, lz0): the
                # reference's build_col_patch / build_row_patch (clipmap.rs:125-170).
                lxs = (lx0 + np.arange(w)) % res
                lzs = (lz0 + np.arange(h)) % res
                uploads.append(dict(level=i, x=x, y=y, w=w, h=h, lx0=lx0, lz0=lz0,
                                    data=texel_heights(height_fn, ox, oz, cell, lxs, lzs)))

            if do_frame and q["moved"]:
                dx, dz = q["dx"], q["dz"]
                if q["first"] or abs(dx) >= res or abs(dz) >= res:
                    off_x = off_z = 0
                    up(0, 0, 0, 0, res, res)
                else:
                    off_x = (off_x + dx) % res
                    off_z = (off_z + dz) % res
                    along_z = (res - off_z) % res if self.wrap_fix else 0
                    along_x = (res - off_x) % res if self.wrap_fix else 0
                    if dx != 0:
                        cols = abs(dx)
                        lx0 = res - cols if dx > 0 else 0
                        sx0 = (lx0 + off_x) % res
                        if sx0 + cols <= res:
                            up(sx0, 0, lx0, along_z, cols, res)
                        else:
                            a = res - sx0
                            up(sx0, 0, lx0, along_z, a, res)
                            up(0, 0, lx0 + a, along_z, cols - a, res)
                    if dz != 0:
                        rows = abs(dz)
                        lz0 = res - rows if dz > 0 else 0
                        sy0 = (lz0 + off_z) % res
                        if sy0 + rows <= res:
                            up(0, sy0, along_x, lz0, res, rows)
                        else:
                            a = res - sy0
                            up(0, sy0, along_x, lz0, res, a)
                            up(0, 0, along_x, lz0 + a, res, rows - a)
                self.tex_offset[i] = (off_x, off_z)
                self.last_cell[i] = (q["cx"], q["cz"])
            cx, cz = self.last_cell[i]
            if cx == I32_MIN:
                params["level"].append((ox, oz, cell))
            else:
                params["level"].append((F(cx) * cell, F(cz) * cell, cell))
            params["offset"].append(self.tex_offset[i])
        return params, uploads


# ---------------------------------------------------------------------------
# The seeded synthetic world's surface height (binary64 value noise), numpy.
# ---------------------------------------------------------------------------
def _mix32(h):
    h = h ^ (h >> np.uint32(16))
    h = h * np.uint32(0x85EBCA6B)
    h = h ^ (h >> np.uint32(13))
    h = h * np.uint32(0xC2B2AE35)
    return h ^ (h >> np.uint32(16))


def _lattice(x, z, salt):
    x = np.asarray(x, np.int64).astype(np.uint32)
    z = np.asarray(z, np.int64).astype(np.uint32)
    with np.errstate(over="ignore"):
        h = _mix32(x * np.uint32(0x9E3779B1) ^ _mix32(z * np.uint32(0x85EBCA77) ^ np.uint32(salt)))
    return (h >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


def _value_noise(x, z, salt):
    fx, fz = np.floor(x), np.floor(z)
    ix, iz = fx.astype(np.int64), fz.astype(np.int64)
    tx, tz = x - fx, z - fz
    ux = tx * tx * (3.0 - 2.0 * tx)
    uz = tz * tz * (3.0 - 2.0 * tz)
    a, b = _lattice(ix, iz, salt), _lattice(ix + 1, iz, salt)
    c, d = _lattice(ix, iz + 1, salt), _lattice(ix + 1, iz + 1, salt)
    x0 = a + (b - a) * ux
    x1 = c + (d - c) * ux
    return x0 + (x1 - x0) * uz


def ground_height_np(seed: int, base_y: int, amplitude: int, wavelength0: int, wx, wz) -> np.ndarray:
    """Surface voxel y of columns (wx, wz): two octaves of value noise at
    wavelengths l0 and l0/2, weights 1 and 1/2."""
    wx = np.asarray(wx, np.int64).astype(np.float64)
    wz = np.asarray(wz, np.int64).astype(np.float64)
    l0 = float(wavelength0 if wavelength0 > 0 else 64)
    s0 = (seed * 0x9E3779B9 + 0x1234) & 0xFFFFFFFF
    s1 = (seed * 0x9E3779B9 + 0xBEEF) & 0xFFFFFFFF
    o0 = _value_noise(wx / l0, wz / l0, s0)
    o1 = _value_noise(wx / (0.5 * l0), wz / (0.5 * l0), s1)
    n = (o0 + 0.5 * o1) / 1.5
    return (base_y + np.floor(amplitude * (2.0 * n - 1.0) + 0.5).astype(np.int64)).astype(np.int32)
