"""MI355X particle hydraulic erosion: Python host over the ``libshy.so`` C ABI.

The reference's interface is the static C++ API of ``World`` and ``Drop``
(source/world.h:17-36, source/water.h:12-39).  This module mirrors it:

* :class:`Params` -- the static parameters (water.h:43-50, world.h:42-44).
* :class:`World` -- one terrain on the GPU.  ``World.erode(cycles)`` is
  ``World::erode`` (world.h:54-88); ``World.cascade(pos)`` is
  ``World::cascade`` (world.h:90-168); ``World.map`` exposes the post-erode
  read accessors of ``quad::map`` (cellpool.h:413-447).
* :class:`Drop` -- ``Drop(pos)`` / ``Drop.descend()`` (water.h:14, 58-156).

Every computation runs in the HIP library; if ``libshy.so`` is missing the
import of :func:`lib` raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshy.so")

SHY_OK, SHY_ERR_ARG, SHY_ERR_HIP, SHY_ERR_NODEV, SHY_ERR_NONFINITE, SHY_ERR_STATE, SHY_ERR_IO = \
    0, -1, -2, -3, -4, -5, -6
FIELDS = ("height", "discharge", "momentumx", "momentumy", "discharge_track",
          "momentumx_track", "momentumy_track", "rootdensity", "erf_discharge")
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}


class ShyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"shy error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """The reference's statics, same names and defaults."""

    _fields_ = [(n, C.c_float) for n in (
        "evapRate", "depositionRate", "minVol", "maxAge", "entrainment", "gravity",
        "momentumTransfer", "lrate", "maxdiff", "settling")]

    def __init__(self, **kw):
        super().__init__()
        lib().shy_default_params(C.byref(self))
        for k, v in kw.items():
            if k not in dict(self._fields_):
                raise TypeError(f"unknown parameter {k}")
            setattr(self, k, v)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


ENSEMBLE_SIDES = (128, 256, 384, 512, 768, 1024)


def ensemble_spec(n_terrains: int = 64, seed: int = 1) -> list:
    """The batched-ensemble workload: ``n_terrains`` terrains with sides drawn
    uniformly from ENSEMBLE_SIDES, each with its own fBm seed (also its
    World::SEED), evapRate uniform in [0.0005, 0.002], depositionRate uniform
    in [0.05, 0.2] and one drop per 16 cells.  Plain dicts {nx, ny, drops,
    seed, evapRate, depositionRate}; the draws come from numpy's PCG64 stream
    of ``seed`` in that order, terrain by terrain."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_terrains):
        s = int(rng.choice(ENSEMBLE_SIDES))
        e = float(rng.uniform(0.0005, 0.002))
        dr = float(rng.uniform(0.05, 0.2))
        out.append(dict(nx=s, ny=s, drops=s * s // 16, seed=int(rng.integers(1, 2**62)),
                        evapRate=e, depositionRate=dr))
    return out


class MapHeader(C.Structure):
    """shy_map_header of include/shy.h (72 bytes)."""
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("nx", C.c_int32),
                ("ny", C.c_int32), ("cycle", C.c_uint32), ("seed", C.c_uint64),
                ("params", Params)]


MAP_MAGIC, MAP_VERSION = b"SHYMAP01", 1


def write_map(path: str, cells, params: Params | None = None, seed: int = 1, cycle: int = 0):
    """Write a map dump (the format shy_save_map writes) from AoS cells (nx, ny, 8)."""
    a = np.ascontiguousarray(cells, np.float32)
    if a.ndim != 3 or a.shape[2] != 8:
        raise ValueError("cells must be (nx, ny, 8)")
    hd = MapHeader(MAP_MAGIC, MAP_VERSION, a.shape[0], a.shape[1], cycle, seed,
                   params if params is not None else Params())
    with open(path, "wb") as f:
        f.write(bytes(hd))
        f.write(a.tobytes())


def read_map(path: str):
    """(header, cells (nx, ny, 8) float32) of a map dump; no GPU needed."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < C.sizeof(MapHeader):
        raise ValueError(f"{path}: too short for a map dump")
    hd = MapHeader.from_buffer_copy(raw[:C.sizeof(MapHeader)])
    if hd.magic != MAP_MAGIC or hd.version != MAP_VERSION:
        raise ValueError(f"{path}: not a map dump")
    n = hd.nx * hd.ny * 8
    body = np.frombuffer(raw, np.float32, count=n, offset=C.sizeof(MapHeader))
    return hd, body.reshape(hd.nx, hd.ny, 8).copy()


class PlantParams(C.Structure):
    """shy_plant_params: the Plant statics of source/vegetation.h:37-41."""
    _fields_ = [(n, C.c_float) for n in ("maxSize", "growRate", "maxSteep", "maxDischarge",
                                         "maxTreeHeight")]

    def __init__(self, **kw):
        super().__init__(1.5, 0.05, 0.8, 0.3, 0.8)
        for k, v in kw.items():
            setattr(self, k, v)


class DropStep(C.Structure):
    _fields_ = [("age", C.c_int32), ("status", C.c_int32), ("pos", C.c_float * 2),
                ("speed", C.c_float * 2), ("volume", C.c_float), ("sediment", C.c_float),
                ("erosion", C.c_float), ("h2", C.c_float)]


STEP_DTYPE = np.dtype([("age", "<i4"), ("status", "<i4"), ("pos", "<f4", (2,)),
                       ("speed", "<f4", (2,)), ("volume", "<f4"), ("sediment", "<f4"),
                       ("erosion", "<f4"), ("h2", "<f4")])


class DropState(C.Structure):
    _fields_ = [("age", C.c_int32), ("pos", C.c_float * 2), ("speed", C.c_float * 2),
                ("volume", C.c_float), ("sediment", C.c_float)]


class TerrainDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("drops", C.c_int), ("seed", C.c_uint64),
                ("params", Params)]


class Tuning(C.Structure):
    _fields_ = [("walk_block", C.c_int), ("walk_grid", C.c_int), ("tile", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "steps", "walked", "skipped", "exits", "ends_age", "ends_vol", "export40", "gain40",
        "walk_dh", "casc_moves")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load ``libshy.so``; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch's ROCm build ships its own HIP runtime under the same soname
        # (libamdhip64.so.7); the first one loaded serves the whole process,
        # and torch does not run on a newer one loaded before it.  Load
        # torch's first when it is installed, so both share it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                               "or python -m simplehydrology_amd.build")
        L = C.CDLL(LIB_PATH)
        vp, fp, i32, u32, u64 = C.c_void_p, C.POINTER(C.c_float), C.c_int, C.c_uint32, C.c_uint64
        sig = {
            "shy_default_params": (None, [vp]),
            "shy_device_count": (i32, []),
            "shy_create": (i32, [vp, i32, i32, u64, C.POINTER(vp)]),
            "shy_destroy": (i32, [vp]),
            "shy_last_error": (C.c_char_p, [vp]),
            "shy_set_params": (i32, [vp, vp]),
            "shy_get_params": (i32, [vp, vp]),
            "shy_size": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
            "shy_upload_cells": (i32, [vp, vp]),
            "shy_download_cells": (i32, [vp, vp]),
            "shy_upload_field": (i32, [vp, i32, vp]),
            "shy_download_field": (i32, [vp, i32, vp]),
            "shy_generate_fbm": (i32, [vp, u64, i32, C.c_float, C.c_float, C.c_float]),
            "shy_generate_reference": (i32, [vp, i32, i32]),
            "shy_download_textures": (i32, [vp, vp, vp, vp]),
            "shy_download_vertices": (i32, [vp, vp]),
            "shy_save_map": (i32, [vp, C.c_char_p]),
            "shy_default_plant_params": (None, [vp]),
            "shy_vegetation_enable": (i32, [vp, i32, vp]),
            "shy_vegetation_grow": (i32, [vp]),
            "shy_vegetation_count": (i32, [vp, C.POINTER(i32), C.POINTER(C.c_int64)]),
            "shy_vegetation_download": (i32, [vp, vp]),
            "shy_vegetation_upload": (i32, [vp, vp, i32]),
            "shy_upload_cells_tiled": (i32, [vp, i32, vp]),
            "shy_download_cells_tiled": (i32, [vp, i32, vp]),
            "shy_load_map": (i32, [vp, C.c_char_p]),
            "shy_terrain_generate_reference": (i32, [vp, i32, i32, i32]),
            "shy_erode": (i32, [vp, i32]),
            "shy_erode_cycles": (i32, [vp, i32, i32, i32]),
            "shy_trace_drop": (i32, [vp, C.c_float, C.c_float, i32, vp, C.POINTER(i32)]),
            "shy_cascade": (i32, [vp, C.c_float, C.c_float]),
            "shy_descend": (i32, [vp, vp, C.POINTER(i32)]),
            "shy_set_seed": (i32, [vp, u64, u32]),
            "shy_get_cycle": (i32, [vp, C.POINTER(u32)]),
            "shy_get_stats": (i32, [vp, vp]),
            "shy_reset_stats": (i32, [vp]),
            "shy_synchronize": (i32, [vp]),
            "shy_set_walk_block": (i32, [vp, i32]),
            "shy_set_drop_order": (i32, [vp, i32]),
            "shy_time_cycles": (i32, [vp, i32, i32, fp]),
            "shy_ensemble_create": (i32, [vp, i32, C.POINTER(vp)]),
            "shy_ensemble_count": (i32, [vp, C.POINTER(i32)]),
            "shy_terrain_size": (i32, [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
            "shy_terrain_set_params": (i32, [vp, i32, vp]),
            "shy_terrain_set_drops": (i32, [vp, i32, i32]),
            "shy_terrain_upload_field": (i32, [vp, i32, i32, vp]),
            "shy_terrain_download_field": (i32, [vp, i32, i32, vp]),
            "shy_terrain_generate_fbm": (i32, [vp, i32, u64, i32, C.c_float, C.c_float,
                                              C.c_float]),
            "shy_terrain_steps": (i32, [vp, vp]),
            "shy_terrain_drop_log": (i32, [vp, i32, vp, vp, i32, C.POINTER(C.c_int)]),
            "shy_ensemble_erode": (i32, [vp, i32, i32]),
            "shy_set_cycle": (i32, [vp, u32]),
            "shy_set_tuning": (i32, [vp, vp]),
            "shy_get_tuning": (i32, [vp, vp]),
            "shy_autotune": (i32, [vp, i32, vp, fp]),
            "shy_terrain_export": (i32, [vp, i32, vp]),
            "shy_terrain_import_flow": (i32, [vp, i32, vp]),
            "shy_download_normals": (i32, [vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_count() -> int:
    return lib().shy_device_count()


def _torch_sync():
    """Device buffers handed to export/import may be produced or consumed by
    work queued on torch's stream; the library orders only its own stream."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        torch.cuda.synchronize()


def _cycle_times(ms) -> dict:
    """Slots of shy_time_cycles: total, then the five kernel phases."""
    return {"ms_total": ms[0], "ms_pack": ms[1], "ms_count_walk": ms[2],
            "ms_partition": ms[3], "ms_replay": ms[4], "ms_reduce_cascade": ms[5]}


def _drop_log(ctx, terrain: int):
    """(steps, ends) of the last cycle's drops of one terrain (see
    shy_terrain_drop_log; ends: 0 maxAge, 1 minVol, 2 left the map, 3 not
    walked)."""
    nx, ny, drops = C.c_int(), C.c_int(), C.c_int()
    _check(ctx, lib().shy_terrain_size(ctx, terrain, C.byref(nx), C.byref(ny), C.byref(drops)))
    cap = max(drops.value, 0)
    steps = np.zeros(cap, np.int32)
    ends = np.zeros(cap, np.int32)
    n = C.c_int(0)
    _check(ctx, lib().shy_terrain_drop_log(ctx, terrain, steps.ctypes.data, ends.ctypes.data, cap,
                                           C.byref(n)))
    return steps[:n.value], ends[:n.value]


def _check(ctx, rc: int):
    if rc != SHY_OK:
        msg = lib().shy_last_error(ctx)
        raise ShyError(rc, msg.decode() if msg else "")


def _f32(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class MapView:
    """Read accessors of ``quad::map`` (source/cellpool.h:413-447) over a
    snapshot downloaded from the device after ``erode``."""

    def __init__(self, world: "World"):
        self._w = world

    def oob(self, p) -> bool:
        x, y = int(p[0]), int(p[1])
        return x < 0 or y < 0 or x >= self._w.nx or y >= self._w.ny

    def height(self, p) -> float:
        if self.oob(p):
            return 0.0
        return float(self._w.field("height")[int(p[0]), int(p[1])])

    def discharge(self, p) -> float:
        """``erf(0.4 * discharge)``, the value node::discharge returns."""
        if self.oob(p):
            return 0.0
        return float(self._w.field("erf_discharge")[int(p[0]), int(p[1])])

    def normal(self, p):
        """``quad::_normal`` (cellpool.h:181-204), evaluated on the device."""
        return self._w.normals()[int(p[0]), int(p[1])]

    def getCell(self, p):
        if self.oob(p):
            return None
        c = self._w.cells()
        return c[int(p[0]), int(p[1])]


class World:
    """One terrain resident in HBM; mirrors the reference ``World``."""

    def __init__(self, nx: int, ny: int | None = None, seed: int = 1,
                 params: Params | None = None):
        ny = nx if ny is None else ny
        self.nx, self.ny = int(nx), int(ny)
        self._ctx = C.c_void_p()
        p = params if params is not None else Params()
        _check(None, lib().shy_create(C.byref(p), self.nx, self.ny, seed, C.byref(self._ctx)))
        self.map = MapView(self)

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if self._ctx:
            lib().shy_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- parameters and seed -----------------------------------------------
    @property
    def params(self) -> Params:
        p = Params()
        _check(self._ctx, lib().shy_get_params(self._ctx, C.byref(p)))
        return p

    @params.setter
    def params(self, p: Params):
        _check(self._ctx, lib().shy_set_params(self._ctx, C.byref(p)))

    def set_seed(self, seed: int, cycle: int = 0):
        _check(self._ctx, lib().shy_set_seed(self._ctx, seed, cycle))

    @property
    def cycle(self) -> int:
        c = C.c_uint32()
        _check(self._ctx, lib().shy_get_cycle(self._ctx, C.byref(c)))
        return c.value

    # -- map transfer --------------------------------------------------------
    def upload(self, field: str, values):
        a = _f32(values, (self.nx, self.ny))
        _check(self._ctx, lib().shy_upload_field(self._ctx, FIELD_ID[field], a.ctypes.data))

    def field(self, field: str) -> np.ndarray:
        out = np.empty((self.nx, self.ny), np.float32)
        _check(self._ctx, lib().shy_download_field(self._ctx, FIELD_ID[field], out.ctypes.data))
        return out

    def normals(self) -> np.ndarray:
        out = np.empty((self.nx, self.ny, 3), np.float32)
        _check(self._ctx, lib().shy_download_normals(self._ctx, out.ctypes.data))
        return out

    def textures(self, water_rgb=(0.2, 0.4, 1.0)):
        """(dischargeMap, momentumMap) of the reference's frame loop, (nx, ny, 4) each."""
        dis = np.empty((self.nx, self.ny, 4), np.float32)
        mom = np.empty((self.nx, self.ny, 4), np.float32)
        rgb = np.asarray(water_rgb, np.float32)
        _check(self._ctx, lib().shy_download_textures(self._ctx, rgb.ctypes.data, dis.ctypes.data,
                                                       mom.ctypes.data))
        return dis, mom

    def vertices(self) -> np.ndarray:
        """updatenode's vertex data: (nx, ny, 4, 3) = position, normal, tangent, bitangent."""
        out = np.empty((self.nx, self.ny, 4, 3), np.float32)
        _check(self._ctx, lib().shy_download_vertices(self._ctx, out.ctypes.data))
        return out

    def save(self, path: str):
        """Map dump (header + AoS cells, see shy.h); restores with ``load``."""
        _check(self._ctx, lib().shy_save_map(self._ctx, os.fsencode(path)))

    def load(self, path: str):
        _check(self._ctx, lib().shy_load_map(self._ctx, os.fsencode(path)))

    def cells_tiled(self, tileres: int = 512) -> np.ndarray:
        """Cells in the reference's multi-tile pool order: (mapsize^2, tileres, tileres, 8)."""
        m = self.nx // tileres
        out = np.empty((m * m, tileres, tileres, 8), np.float32)
        _check(self._ctx, lib().shy_download_cells_tiled(self._ctx, tileres, out.ctypes.data))
        return out

    def upload_cells_tiled(self, pool, tileres: int = 512):
        m = self.nx // tileres
        a = _f32(pool, (m * m, tileres, tileres, 8))
        _check(self._ctx, lib().shy_upload_cells_tiled(self._ctx, tileres, a.ctypes.data))

    # -- vegetation (vegetation.h:44-186) ----------------------------------
    def vegetation_enable(self, capacity: int = 1 << 20, params: "PlantParams | None" = None):
        p = params if params is not None else PlantParams()
        _check(self._ctx, lib().shy_vegetation_enable(self._ctx, capacity, C.byref(p)))

    def vegetation_grow(self, steps: int = 1):
        """``Vegetation::grow()`` ``steps`` times."""
        for _ in range(steps):
            _check(self._ctx, lib().shy_vegetation_grow(self._ctx))

    def plants(self) -> np.ndarray:
        """The plant list in order: (n, 3) = x, y, size."""
        n, dropped = C.c_int(), C.c_int64()
        _check(self._ctx, lib().shy_vegetation_count(self._ctx, C.byref(n), C.byref(dropped)))
        out = np.empty((n.value, 3), np.float32)
        if n.value:
            _check(self._ctx, lib().shy_vegetation_download(self._ctx, out.ctypes.data))
        return out

    def set_plants(self, xys):
        a = np.ascontiguousarray(xys, np.float32).reshape(-1, 3)
        _check(self._ctx, lib().shy_vegetation_upload(self._ctx, a.ctypes.data, a.shape[0]))

    def upload_cells(self, cells):
        a = _f32(cells, (self.nx, self.ny, 8))
        _check(self._ctx, lib().shy_upload_cells(self._ctx, a.ctypes.data))

    def cells(self) -> np.ndarray:
        out = np.empty((self.nx, self.ny, 8), np.float32)
        _check(self._ctx, lib().shy_download_cells(self._ctx, out.ctypes.data))
        return out

    def generate_fbm(self, seed: int, octaves: int = 8, base_freq: float = 0.0,
                     lacunarity: float = 2.0, gain: float = 0.5):
        _check(self._ctx, lib().shy_generate_fbm(self._ctx, seed, octaves, base_freq,
                                                  lacunarity, gain))

    def generate_reference(self, seed: int, tileres: int = 512):
        """``quad::map::init``'s OpenSimplex2 terrain (cellpool.h:349-409)."""
        _check(self._ctx, lib().shy_generate_reference(self._ctx, int(seed), int(tileres)))

    # -- the hot path ----------------------------------------------------------
    def erode(self, cycles: int):
        """``World::erode(cycles)``: one cycle of ``cycles`` drop spawns."""
        _check(self._ctx, lib().shy_erode(self._ctx, cycles))

    def erode_cycles(self, n_cycles: int, drops: int, graph: bool = True):
        _check(self._ctx, lib().shy_erode_cycles(self._ctx, n_cycles, drops, 1 if graph else 0))

    def cascade(self, pos):
        """``World::cascade(pos)`` on the live map."""
        _check(self._ctx, lib().shy_cascade(self._ctx, float(pos[0]), float(pos[1])))

    def trace(self, x: float, y: float, max_steps: int = 600) -> np.ndarray:
        """Per-step states of one drop on the frozen current map."""
        buf = np.zeros(max_steps, STEP_DTYPE)
        n = C.c_int(0)
        _check(self._ctx, lib().shy_trace_drop(self._ctx, x, y, max_steps, buf.ctypes.data,
                                               C.byref(n)))
        return buf[:min(n.value, max_steps)]

    def drop(self, pos) -> "Drop":
        return Drop(self, pos)

    def stats(self) -> dict:
        s = Stats()
        _check(self._ctx, lib().shy_get_stats(self._ctx, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        _check(self._ctx, lib().shy_reset_stats(self._ctx))

    def synchronize(self):
        _check(self._ctx, lib().shy_synchronize(self._ctx))

    def set_walk_block(self, threads: int):
        _check(self._ctx, lib().shy_set_walk_block(self._ctx, threads))

    def set_drop_order(self, on):
        """Spatial drop order of the count walk: True, False or None (automatic,
        the default: on from 262144 drops per cycle); bit-identical results."""
        _check(self._ctx, lib().shy_set_drop_order(self._ctx, -1 if on is None else int(bool(on))))

    def tuning(self) -> dict:
        t = Tuning()
        _check(self._ctx, lib().shy_get_tuning(self._ctx, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, walk_block=256, walk_grid=0, tile=0):
        t = Tuning(walk_block, walk_grid, tile)
        _check(self._ctx, lib().shy_set_tuning(self._ctx, C.byref(t)))

    def autotune(self, cycles: int = 3):
        """Pick the fastest launch shape (results are identical under all)."""
        t, ms = Tuning(), C.c_float()
        _check(self._ctx, lib().shy_autotune(self._ctx, cycles, C.byref(t), C.byref(ms)))
        return t.as_dict(), ms.value

    def export(self, dst_ptr: int):
        """(h, d, mx, my) into a device buffer (4 floats per cell)."""
        _torch_sync()
        _check(self._ctx, lib().shy_terrain_export(self._ctx, 0, dst_ptr))

    def import_flow(self, src_ptr: int):
        """(d, mx, my) from a device buffer (3 floats per cell)."""
        _torch_sync()
        _check(self._ctx, lib().shy_terrain_import_flow(self._ctx, 0, src_ptr))

    def drop_log(self):
        """(steps, ends) of the last cycle's walk, one entry per drop."""
        return _drop_log(self._ctx, 0)

    def time_cycles(self, cycles: int, drops: int) -> dict:
        """Run ``cycles`` erode calls with HIP events between kernels on the
        context's stream; returns milliseconds per kernel, summed."""
        ms = (C.c_float * 6)()
        _check(self._ctx, lib().shy_time_cycles(self._ctx, cycles, drops, ms))
        return _cycle_times(ms)


class Drop:
    """``Drop`` (source/water.h:12-39): state on the host, steps on the GPU."""

    def __init__(self, world: World, pos):
        self.world = world
        self.state = DropState()
        self.state.age = 0
        self.state.pos[0], self.state.pos[1] = float(pos[0]), float(pos[1])
        self.state.speed[0] = self.state.speed[1] = 0.0
        self.state.volume = 1.0
        self.state.sediment = 0.0

    age = property(lambda self: self.state.age)
    pos = property(lambda self: (self.state.pos[0], self.state.pos[1]))
    speed = property(lambda self: (self.state.speed[0], self.state.speed[1]))
    volume = property(lambda self: self.state.volume)
    sediment = property(lambda self: self.state.sediment)

    def descend(self) -> bool:
        """One ``Drop::descend()`` call on the live map."""
        alive = C.c_int(0)
        _check(self.world._ctx, lib().shy_descend(self.world._ctx, C.byref(self.state),
                                                  C.byref(alive)))
        return bool(alive.value)


class Ensemble(World):
    """Many terrains of uneven sizes and parameter sets in one context: every
    cycle runs all of them in one launch per kernel (hipGraph-replayable)."""

    def __init__(self, terrains):
        """terrains: iterable of dicts {nx, ny, drops, seed, params (Params)}."""
        ts = list(terrains)
        arr = (TerrainDesc * len(ts))()
        for i, t in enumerate(ts):
            arr[i].nx, arr[i].ny = int(t["nx"]), int(t.get("ny", t["nx"]))
            arr[i].drops, arr[i].seed = int(t["drops"]), int(t["seed"])
            arr[i].params = t.get("params") or Params()
        self._ctx = C.c_void_p()
        _check(None, lib().shy_ensemble_create(arr, len(ts), C.byref(self._ctx)))
        self.nter = len(ts)
        self.sizes = [(arr[i].nx, arr[i].ny) for i in range(len(ts))]
        self.nx, self.ny = self.sizes[0]
        self.map = None

    def generate_fbm(self, terrain: int, seed: int, octaves: int = 8, base_freq: float = 0.0,
                     lacunarity: float = 2.0, gain: float = 0.5):
        _check(self._ctx, lib().shy_terrain_generate_fbm(self._ctx, terrain, seed, octaves,
                                                          base_freq, lacunarity, gain))

    def generate_reference(self, terrain: int, seed: int, tileres: int = 512):
        _check(self._ctx, lib().shy_terrain_generate_reference(self._ctx, terrain, int(seed),
                                                                int(tileres)))

    def field(self, terrain: int, field: str) -> np.ndarray:
        nx, ny = self.sizes[terrain]
        out = np.empty((nx, ny), np.float32)
        _check(self._ctx, lib().shy_terrain_download_field(self._ctx, terrain, FIELD_ID[field],
                                                            out.ctypes.data))
        return out

    def upload(self, terrain: int, field: str, values):
        a = _f32(values, self.sizes[terrain])
        _check(self._ctx, lib().shy_terrain_upload_field(self._ctx, terrain, FIELD_ID[field],
                                                          a.ctypes.data))

    def erode(self, cycles: int = 1, graph: bool = True):
        _check(self._ctx, lib().shy_ensemble_erode(self._ctx, cycles, 1 if graph else 0))

    def export(self, terrain: int, dst_ptr: int):
        """(h, d, mx, my) of one terrain into a device buffer (4 floats/cell)."""
        _torch_sync()
        _check(self._ctx, lib().shy_terrain_export(self._ctx, terrain, dst_ptr))

    def import_flow(self, terrain: int, src_ptr: int):
        """(d, mx, my) of one terrain from a device buffer (3 floats/cell)."""
        _torch_sync()
        _check(self._ctx, lib().shy_terrain_import_flow(self._ctx, terrain, src_ptr))

    def drop_log(self, terrain: int):
        """(steps, ends) of the last cycle's walk of one terrain's drops."""
        return _drop_log(self._ctx, terrain)

    def terrain_steps(self) -> np.ndarray:
        out = np.zeros(self.nter, np.int64)
        _check(self._ctx, lib().shy_terrain_steps(self._ctx, out.ctypes.data))
        return out

    def time_cycles(self, cycles: int) -> dict:
        ms = (C.c_float * 6)()
        _check(self._ctx, lib().shy_time_cycles(self._ctx, cycles, -1, ms))
        return _cycle_times(ms)
