"""Python mirror of the renderer API over the C ABI (tests and the bench use it).

Method names follow ``src/renderers/GaussianSplatRenderer.ts`` (setTileSize
:301-317, getTileCount/getTileSize :1268-1274, readTileCounts :1278-1310,
rebin :1104-1266) and ``src/SceneSyncer.ts`` (setSplatData :25-43).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .scene import Splats


class GaussianSplatRenderer:
    """Owns one gs3d context (device memory + one HIP stream) on `device`."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = C.c_void_p()
        rc = self._lib.gs3d_create(device, C.byref(h))
        if rc != L.GS3D_OK:
            raise L.Gs3dError(rc, f"gs3d_create(device={device}) failed")
        self._h = h
        self.splat_count = 0

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != L.GS3D_OK:
            raise L.Gs3dError(rc, self._lib.gs3d_last_error(self._h).decode())

    def close(self) -> None:
        if self._h:
            self._lib.gs3d_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        self._check(self._lib.gs3d_sync(self._h))

    def device_limits(self) -> dict:
        out = np.zeros(6, np.uint32)
        self._check(self._lib.gs3d_device_limits(self._h, L.ptr(out)))
        keys = ("maxTotalThreads", "maxSizeX", "maxSizeY", "maxSizeZ", "maxDispatch",
                "maxSharedMemory")
        return dict(zip(keys, (int(v) for v in out)))

    # -- scene --------------------------------------------------------------
    def set_splat_data(self, s: Splats) -> None:
        c = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        pos, sc, rot, col = c(s.positions), c(s.scales), c(s.rotations), c(s.colors)
        dc, rest = c(s.f_dc), c(s.sh_rest)
        self._keep = (pos, sc, rot, col, dc, rest)
        self._check(self._lib.gs3d_set_splats(
            self._h, L.ptr(pos), L.ptr(sc), L.ptr(rot), L.ptr(col),
            L.ptr(dc) if s.sh_degree else None, L.ptr(rest) if s.sh_degree else None,
            s.sh_degree, s.n))
        self._keep = None
        self.splat_count = s.n

    def load_ply(self, path, threads: int = 0) -> L.PlyInfo:
        """Parse a PLY file natively and upload it (gs3d_load_ply)."""
        info = L.PlyInfo()
        self._check(self._lib.gs3d_load_ply(self._h, str(path).encode(), threads, C.byref(info)))
        self.splat_count = int(info.splat_count)
        return info

    # -- renderer API ---------------------------------------------------------
    def set_tile_size(self, x: int, y: int) -> tuple[int, int]:
        out = np.zeros(2, np.uint32)
        self._check(self._lib.gs3d_set_tile_size(self._h, x, y, L.ptr(out)))
        return int(out[0]), int(out[1])

    def get_tile_size(self) -> tuple[int, int]:
        out = np.zeros(2, np.uint32)
        self._check(self._lib.gs3d_get_tile_size(self._h, L.ptr(out)))
        return int(out[0]), int(out[1])

    def set_kernel_variant(self, kernel: str, param: str, value: int) -> None:
        self._check(self._lib.gs3d_set_kernel_variant(self._h, kernel.encode(), param.encode(), value))

    def get_kernel_variant(self, kernel: str, param: str) -> int:
        v = C.c_int()
        self._check(self._lib.gs3d_get_kernel_variant(self._h, kernel.encode(), param.encode(),
                                                      C.byref(v)))
        return int(v.value)

    def get_tile_count(self) -> tuple[int, int]:
        out = np.zeros(2, np.uint32)
        self._check(self._lib.gs3d_get_tile_count(self._h, L.ptr(out)))
        return int(out[0]), int(out[1])

    def rebin(self, camera_uniform: np.ndarray, width: int, height: int, flags: int = 0) -> None:
        cam = np.ascontiguousarray(camera_uniform, np.float32)
        if cam.size != 72:
            raise ValueError("camera uniform must hold 72 floats")
        self._check(self._lib.gs3d_rebin(self._h, L.ptr(cam), width, height, flags))

    def ref_count(self) -> int:
        k = C.c_uint64()
        self._check(self._lib.gs3d_ref_count(self._h, C.byref(k)))
        return int(k.value)

    def read(self, buffer_id: int, dtype, shape=None) -> np.ndarray:
        size = C.c_size_t()
        self._check(self._lib.gs3d_read(self._h, buffer_id, None, 0, C.byref(size)))
        out = np.empty(size.value // np.dtype(dtype).itemsize, dtype)
        self._check(self._lib.gs3d_read(self._h, buffer_id, L.ptr(out), out.nbytes, C.byref(size)))
        return out.reshape(shape) if shape is not None else out

    def read_projected(self) -> np.ndarray:
        return self.read(L.BUF_PROJECTED_SPLATS, np.float32).reshape(-1, 16)

    def read_image(self, width: int, height: int) -> np.ndarray:
        """H x W x 4 float32 (the rgba16float target widened)."""
        return self.read(L.BUF_IMAGE_RGBA32F, np.float32).reshape(height, width, 4)

    def read_tile_counts(self) -> np.ndarray | None:
        """Per-tile counts with the underflow guard; None on failure (readTileCounts)."""
        n = C.c_size_t()
        if self._lib.gs3d_read_tile_counts(self._h, None, 0, C.byref(n)) != L.GS3D_OK:
            return None
        out = np.empty(n.value, np.uint32)
        if self._lib.gs3d_read_tile_counts(self._h, L.ptr(out), out.size, C.byref(n)) != L.GS3D_OK:
            return None
        return out

    # -- batched views ----------------------------------------------------------
    @staticmethod
    def _views(views) -> np.ndarray:
        arr = np.zeros(len(views), dtype=L.VIEW_DTYPE)
        for i, (cam, w, h) in enumerate(views):
            cam = np.asarray(cam, np.float32).reshape(-1)
            if cam.size != 72:
                raise ValueError("camera uniform must hold 72 floats")
            arr[i]["camera"] = cam
            arr[i]["width"] = w
            arr[i]["height"] = h
        return arr

    def rebin_batch(self, views, eager: bool = False, learned: bool = False) -> None:
        """Render many views: `views` is a list of (camera uniform (72 f32), width, height).
        learned: colour with the learned stage when it is enabled (else exact SH)."""
        arr = np.zeros(len(views), dtype=L.VIEW_DTYPE)
        for i, (cam, w, h) in enumerate(views):
            cam = np.asarray(cam, np.float32).reshape(-1)
            if cam.size != 72:
                raise ValueError("camera uniform must hold 72 floats")
            arr[i]["camera"] = cam
            arr[i]["width"] = w
            arr[i]["height"] = h
        flags = (L.BATCH_EAGER if eager else 0) | (L.BATCH_LEARNED if learned else 0)
        self._check(self._lib.gs3d_rebin_batch(self._h, L.ptr(arr), len(views), flags))
        self._batch_views = [(int(w), int(h)) for _, w, h in views]

    def batch_ref_counts(self) -> np.ndarray:
        out = np.zeros(len(self._batch_views), np.uint64)
        self._check(self._lib.gs3d_batch_ref_counts(self._h, L.ptr(out), out.size))
        return out

    def batch_read(self, view: int, buffer_id: int, dtype) -> np.ndarray:
        size = C.c_size_t()
        self._check(self._lib.gs3d_batch_read(self._h, view, buffer_id, None, 0, C.byref(size)))
        out = np.empty(size.value // np.dtype(dtype).itemsize, dtype)
        self._check(self._lib.gs3d_batch_read(self._h, view, buffer_id, L.ptr(out), out.nbytes,
                                              C.byref(size)))
        return out

    # -- learned colour -------------------------------------------------------------
    def learned_init(self, seed: int = 0) -> None:
        self._check(self._lib.gs3d_learned_init(self._h, seed))

    def learned_train(self, epochs: int, lr: float = 1e-3, lr_latent: float = 1e-2) -> np.ndarray:
        """Train; returns the mean squared error of each epoch."""
        loss = np.zeros(max(1, epochs), np.float32)
        self._check(self._lib.gs3d_learned_train(self._h, epochs, lr, lr_latent, L.ptr(loss), loss.size))
        return loss[:epochs]

    def learned_grad(self, first_dir: int):
        gp = np.zeros(L.LEARNED_PARAMS, np.float32)
        gl = np.zeros((self.splat_count, 8), np.float32)
        loss = C.c_float()
        self._check(self._lib.gs3d_learned_grad(self._h, first_dir, L.ptr(gp), L.ptr(gl), C.byref(loss)))
        return float(loss.value), gp, gl

    def learned_get(self):
        p = np.zeros(L.LEARNED_PARAMS, np.float32)
        z = np.zeros((self.splat_count, 8), np.float32)
        z16 = np.zeros((self.splat_count, 8), np.uint16)
        self._check(self._lib.gs3d_learned_get(self._h, L.ptr(p), L.ptr(z), L.ptr(z16)))
        return p, z, z16

    def learned_set(self, params=None, latents=None) -> None:
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        z = None if latents is None else np.ascontiguousarray(latents, np.float32)
        self._check(self._lib.gs3d_learned_set(self._h, L.ptr(p) if p is not None else None,
                                               L.ptr(z) if z is not None else None))

    STATE_ARRAYS = ("params", "params_m", "params_v", "latents", "latents_m", "latents_v")

    def learned_get_state(self) -> dict:
        """The whole training state (a checkpoint): the six arrays and the Adam step."""
        out = {k: np.zeros(L.LEARNED_PARAMS, np.float32) for k in self.STATE_ARRAYS[:3]}
        out.update({k: np.zeros((self.splat_count, 8), np.float32) for k in self.STATE_ARRAYS[3:]})
        st = L.LearnedState(*[L.ptr(out[k]) for k in self.STATE_ARRAYS], 0)
        self._check(self._lib.gs3d_learned_get_state(self._h, C.byref(st)))
        out["step"] = int(st.step)
        return out

    def learned_set_state(self, state: dict) -> None:
        arrays = {k: np.ascontiguousarray(state[k], np.float32) for k in self.STATE_ARRAYS if k in state}
        st = L.LearnedState(*[L.ptr(arrays[k]) if k in arrays else None for k in self.STATE_ARRAYS],
                            int(state.get("step", 0)))
        self._check(self._lib.gs3d_learned_set_state(self._h, C.byref(st)))

    @staticmethod
    def learned_dirs() -> np.ndarray:
        d = np.zeros((L.LEARNED_DIRS, 3), np.float32)
        L.lib().gs3d_learned_dirs(L.ptr(d))
        return d

    def learned_colors(self, eye) -> np.ndarray:
        e = np.asarray(eye, np.float32)
        out = np.zeros((self.splat_count, 3), np.float32)
        self._check(self._lib.gs3d_learned_colors(self._h, L.ptr(e), L.ptr(out)))
        return out

    def learned_validate(self, views, min_psnr_db: float = 45.0):
        arr = self._views(views)
        psnr, en = C.c_float(), C.c_int()
        self._check(self._lib.gs3d_learned_validate(self._h, L.ptr(arr), len(views), min_psnr_db,
                                                    C.byref(psnr), C.byref(en)))
        self._batch_views = [(int(w), int(h)) for _, w, h in views]
        return float(psnr.value), bool(en.value)

    def learned_enable(self, on: bool) -> None:
        self._check(self._lib.gs3d_learned_enable(self._h, 1 if on else 0))

    # -- display: heat map + composite -----------------------------------------
    def set_debug_mode(self, mode: int, debug_ref: int = 256) -> None:
        """setDebugMode (GaussianSplatRenderer.ts:319-323); debug_ref = count drawn full red."""
        self._check(self._lib.gs3d_set_debug_mode(self._h, mode, debug_ref))

    def get_debug_mode(self) -> tuple[int, int]:
        m, r = C.c_uint32(), C.c_uint32()
        self._check(self._lib.gs3d_get_debug_mode(self._h, C.byref(m), C.byref(r)))
        return int(m.value), int(r.value)

    def composite(self, width: int, height: int, clear=(0.12, 0.12, 0.13, 1.0), framebuffer=None,
                  fmt: int = L.FORMAT_RGBA8UNORM) -> np.ndarray:
        """Composite the last frame over a framebuffer; returns it as H x W x 4 uint8.

        clear: RGBA floats the framebuffer is cleared to first, or None to blend
        onto `framebuffer` (H x W x 4 uint8, or None for the resident one)."""
        out = (np.zeros((height, width, 4), np.uint8) if framebuffer is None
               else np.ascontiguousarray(framebuffer, np.uint8).copy())
        cl = None if clear is None else np.asarray(clear, np.float32)
        if cl is None and framebuffer is None:
            self._check(self._lib.gs3d_composite(self._h, fmt, None, None, 0))
            return self.read(L.BUF_FRAMEBUFFER, np.uint8).reshape(height, width, 4)
        self._check(self._lib.gs3d_composite(self._h, fmt, L.ptr(cl) if cl is not None else None,
                                             L.ptr(out), out.nbytes))
        return out

    def composite_resident(self, fmt: int = L.FORMAT_RGBA8UNORM) -> None:
        """Composite onto the resident framebuffer, no host copies (asynchronous)."""
        self._check(self._lib.gs3d_composite(self._h, fmt, None, None, 0))

    def stage_times_us(self) -> dict:
        out = np.zeros(len(L.STAGES), np.float32)
        self._check(self._lib.gs3d_stage_times_us(self._h, L.ptr(out), out.size))
        return {k: float(v) for k, v in zip(L.STAGES, out)}

    def profiler_reset(self) -> None:
        self._check(self._lib.gs3d_profiler_reset(self._h))

    def set_profile_mask(self, stages) -> None:
        """Profile only the named stages (labels of _lib.STAGES)."""
        mask = 0
        for name in stages:
            mask |= 1 << L.STAGES.index(name)
        self._check(self._lib.gs3d_set_profile_mask(self._h, mask))

    def stage_times_avg_us(self) -> tuple[dict, int]:
        out = np.zeros(len(L.STAGES), np.float32)
        frames = C.c_uint64()
        self._check(self._lib.gs3d_stage_times_avg_us(self._h, L.ptr(out), out.size,
                                                      C.byref(frames)))
        return {k: float(v) for k, v in zip(L.STAGES, out)}, int(frames.value)

    def sort_passes(self) -> list[int]:
        n = C.c_uint32()
        out = np.zeros(16, np.uint32)
        self._check(self._lib.gs3d_sort_passes(self._h, C.byref(n), L.ptr(out)))
        return [int(v) for v in out[: n.value]]

    # -- stage entry points (each reference pass fed caller data) -------------
    def stage_preprocess(self, camera_uniform, tiles_x, tiles_y, inv_tile_x, inv_tile_y):
        n = self.splat_count
        proj = np.zeros((n, 16), np.float32)
        counts = np.zeros(n, np.uint32)
        cam = np.ascontiguousarray(camera_uniform, np.float32)
        self._check(self._lib.gs3d_stage_preprocess(self._h, L.ptr(cam), tiles_x, tiles_y,
                                                    inv_tile_x, inv_tile_y, L.ptr(proj),
                                                    L.ptr(counts)))
        return proj, counts

    def stage_scan(self, counts: np.ndarray):
        counts = np.ascontiguousarray(counts, np.uint32)
        out = np.zeros(counts.size + 1, np.uint32)
        total = C.c_uint32()
        self._check(self._lib.gs3d_stage_scan(self._h, L.ptr(counts), counts.size, L.ptr(out),
                                              C.byref(total)))
        return out, int(total.value)

    def stage_emit(self, projected: np.ndarray, offsets: np.ndarray, tiles_x: int):
        projected = np.ascontiguousarray(projected, np.float32)
        offsets = np.ascontiguousarray(offsets, np.uint32)
        n = offsets.size - 1
        k = int(offsets[n])
        keys = np.zeros(k, np.uint64)
        vals = np.zeros(k, np.uint32)
        self._check(self._lib.gs3d_stage_emit(self._h, L.ptr(projected), L.ptr(offsets), n, tiles_x,
                                              L.ptr(keys), L.ptr(vals)))
        return keys, vals

    def stage_sort(self, keys: np.ndarray, values: np.ndarray, digit_bits: int = 8):
        keys = np.ascontiguousarray(keys, np.uint64)
        values = np.ascontiguousarray(values, np.uint32)
        ko = np.zeros_like(keys)
        vo = np.zeros_like(values)
        self._check(self._lib.gs3d_stage_sort(self._h, L.ptr(keys), L.ptr(values), keys.size,
                                              digit_bits, L.ptr(ko), L.ptr(vo)))
        return ko, vo

    def stage_identify(self, sorted_keys: np.ndarray, tile_count: int):
        sorted_keys = np.ascontiguousarray(sorted_keys, np.uint64)
        out = np.zeros(tile_count + 1, np.uint32)
        self._check(self._lib.gs3d_stage_identify(self._h, L.ptr(sorted_keys), sorted_keys.size,
                                                  tile_count, L.ptr(out)))
        return out

    def stage_rasterize(self, projected, sorted_values, tile_offsets, tile, width, height):
        projected = np.ascontiguousarray(projected, np.float32)
        sorted_values = np.ascontiguousarray(sorted_values, np.uint32)
        tile_offsets = np.ascontiguousarray(tile_offsets, np.uint32)
        img = np.zeros((height, width, 4), np.float32)
        self._check(self._lib.gs3d_stage_rasterize(
            self._h, L.ptr(projected), projected.shape[0], L.ptr(sorted_values),
            sorted_values.size, L.ptr(tile_offsets), tile, tile, width, height, L.ptr(img)))
        return img

    def stage_composite(self, target_f16, tile_offsets, tile, width, height, debug=0, debug_ref=256,
                        clear=None, framebuffer=None, fmt: int = L.FORMAT_RGBA8UNORM):
        """Heat map + composite over caller data; returns (framebuffer, debug target or None)."""
        target_f16 = np.ascontiguousarray(target_f16, np.uint16)
        assert target_f16.size == width * height * 4
        to = None if tile_offsets is None else np.ascontiguousarray(tile_offsets, np.uint32)
        fb = (np.zeros((height, width, 4), np.uint8) if framebuffer is None
              else np.ascontiguousarray(framebuffer, np.uint8).copy())
        cl = None if clear is None else np.asarray(clear, np.float32)
        dbg = np.zeros((height, width, 4), np.uint16) if debug else None
        self._check(self._lib.gs3d_stage_composite(
            self._h, L.ptr(target_f16), L.ptr(to) if to is not None else None, tile, tile, width,
            height, debug, debug_ref, fmt, L.ptr(cl) if cl is not None else None, L.ptr(fb),
            L.ptr(dbg) if dbg is not None else None))
        return fb, dbg
