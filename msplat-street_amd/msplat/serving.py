"""Serving novel views of resident scenes with continuous batching.

Trained scenes stay resident in HBM. Requests (scene, camera, resolution) arrive
continuously; a request waits in the pending queue until a worker slot is free and the
HBM budget admits it. Each slot owns a context (its own stream and buffers) and runs in
its own thread; the library releases the GIL during its calls, so the renders of
different slots overlap on the device, and requests that arrive meanwhile are admitted
as slots and memory free up. A free slot renders the oldest request alone
(msplat_hip_render) while the queue is short; once more than `batch_depth` requests are
pending, it renders the oldest together with later pending requests of its scene in one
msplat_hip_render_batch (mixed resolutions, one persistent rasterizer launch), up to
`max_batch_pixels` in all. A batch costs about what its views cost one by one, except
that small views fill the device better together; so batches are kept short (every
request of a batch completes when the batch does) and form only when requests queue. Every render ends with one RGBA8
pack (msplat_pack_rgba8, core/cpu/msplat_cpu.cpp:1723-1734). The plain form renders one
request per msplat_hip_render, first come first served, on one context.

Admission: the resident scenes plus every slot's working set stay within the HBM budget;
a batch's working set is estimated from its record count (views x gaussians), pixels and
intersections (pixels x the scene's measured intersections per pixel). Scenes are
admitted and evicted while the server runs (add_scene, remove_scene): a scene is admitted
when, with it resident, every slot's share still holds the smallest view of every
resident scene, and an evicted scene's pending requests are refused and its memory freed
once no slot renders it.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import threading
import time

import numpy as np

from ._abi import TUNE_CHUNKS, Context, DeviceArray, check, floats, lib, make_camera_struct
from .backend import BatchViews, DeviceGaussians, render_batch

REC_BYTES = 116    # per record (view x gaussian) of a batched frame
PIX_BYTES = 40     # per pixel: image, final T and index, gradient image, loss, RGBA8
ISECT_BYTES = 52   # per intersection: keys, ids (double-buffered), gradient slots
PARAM_BYTES = 236  # per resident SH3 gaussian (59 floats)
MIN_VIEW = (640, 360)  # the smallest view every slot must keep room for


@dataclasses.dataclass
class Request:
    rid: int
    scene: int
    cam: dict
    arrival: float                 # seconds from the start of the run
    launched: float = float("nan")
    done: float = float("nan")
    batch_size: int = 0
    slot: int = -1
    image: np.ndarray | None = None  # RGBA8, kept when the server keeps images

    @property
    def pixels(self) -> int:
        return int(self.cam["W"]) * int(self.cam["H"])


class _Slot:
    def __init__(self, index: int, device: int):
        self.index = index
        self.ctx = Context(device)
        # requests render per tile, alone or batched (the batched rasterizer has no depth
        # chunks): a request's image does not depend on the batch it landed in
        self.ctx.set_tuning(TUNE_CHUNKS, 1)
        self.rgba = None
        self.busy = False

    def rgba_buffer(self, pixels: int) -> DeviceArray:
        """The slot's RGBA8 output buffer, grown on the slot's own thread; the old one is
        freed there explicitly (its last pack completed before the slot's sync), never
        by the garbage collector on another thread."""
        if self.rgba is None or self.rgba.nbytes < 4 * pixels:
            old, self.rgba = self.rgba, DeviceArray(self.ctx, (int(pixels * 1.25) + 1024, 4), np.uint8)
            if old is not None:
                old.free()
        return self.rgba

    def close(self):
        if self.rgba is not None:
            self.ctx.sync()
            self.rgba.free()
            self.rgba = None
        self.ctx.close()


class Server:
    """Continuous-batching renderer over resident scenes (DeviceGaussians)."""

    def __init__(self, scenes: list, slots: int = 2, max_views: int = 16, hbm_budget: float | None = None,
                 background=(0.0, 0.0, 0.0), device: int = 0, keep_images: bool = False,
                 isects_per_pixel: float = 60.0, adapt: bool = True, batch_depth: int | None = None,
                 max_batch_pixels: float = 2.1e6, max_wait_ms: float | None = None):
        self.scenes = scenes
        self.max_views = int(max_views)
        # pending requests beyond which a free slot batches (default: 2 per slot), and the
        # pixels a batch may hold in all (default: about one 1920x1080 view)
        self.batch_depth = 2 * slots if batch_depth is None else int(batch_depth)
        self.max_batch_pixels = float(max_batch_pixels)
        # load shedding: a request still queued this long after its arrival is dropped
        # (to self.shed) instead of rendered, so that under overload the served requests
        # keep their latency bound; None serves everything
        self.max_wait = None if max_wait_ms is None else float(max_wait_ms) / 1e3
        self.bg = tuple(float(v) for v in background)
        self.keep_images = keep_images
        self.scenes = list(scenes)  # None where a scene was evicted (indices stay)
        self.resident = sum(s.nbytes for s in self.scenes)
        if hbm_budget is None:  # the resident scenes plus 90% of what the device has free
            from ._abi import device_memory
            hbm_budget = self.resident + 0.9 * device_memory(device)[0]
        self.budget = float(hbm_budget)
        if self.resident >= self.budget:
            raise ValueError("the resident scenes alone exceed the HBM budget")
        self.nslots = slots
        self.slots = [_Slot(i, device) for i in range(slots)]
        self.share = (self.budget - self.resident) / slots
        self.ipp0 = float(isects_per_pixel)
        self.ipp = [self.ipp0] * len(self.scenes)  # intersections per pixel
        self.inflight = [0] * len(self.scenes)  # batches of each scene being rendered
        self.evicting: set = set()
        self.adapt = adapt  # refine ipp from each rendered frame
        self.pending: list[Request] = []
        self.finished: list[Request] = []
        self.rejected: list[Request] = []
        self.shed: list[Request] = []
        self.cond = threading.Condition()
        self.stop = False
        self.t0 = 0.0
        self.batches = 0
        self.errors: list[str] = []

    # ── admission ────────────────────────────────────────────────────────────
    def working_set(self, reqs) -> float:
        return self._working_set(reqs[0].scene, len(reqs), sum(r.pixels for r in reqs))

    def _working_set(self, scene: int, views: int, px: int) -> float:
        return views * self.scenes[scene].n * REC_BYTES + px * PIX_BYTES + px * self.ipp[scene] * ISECT_BYTES

    def _live(self, scene: int) -> bool:
        return 0 <= scene < len(self.scenes) and self.scenes[scene] is not None and scene not in self.evicting

    def add_scene(self, scene) -> int | None:
        """Admits a scene (DeviceGaussians on this device) while the server runs. It is
        refused (None) when, with it resident, a slot's share of the budget would not hold
        the smallest view (MIN_VIEW) of every resident scene; else its index is returned
        and requests for it are served from then on."""
        need = scene.nbytes
        px = MIN_VIEW[0] * MIN_VIEW[1]
        with self.cond:
            share = (self.budget - self.resident - need) / self.nslots
            self.scenes.append(scene)
            self.ipp.append(self.ipp0)
            self.inflight.append(0)
            idx = len(self.scenes) - 1
            fits = all(share >= self._working_set(i, 1, px) for i in range(len(self.scenes)) if self._live(i))
            if not fits:
                self.scenes[idx] = None
                return None
            self.resident += need
            self.share = share
            return idx

    def remove_scene(self, index: int, free: bool = True):
        """Evicts a scene: its pending requests are refused now, and so are later ones; the
        call returns once no slot renders it, having freed its device memory (free=True)
        and returned it to the slots' shares."""
        with self.cond:
            if not self._live(index):
                raise ValueError(f"scene {index} is not resident")
            self.evicting.add(index)
            keep = []
            for r in self.pending:
                (self.rejected if r.scene == index else keep).append(r)
            self.pending = keep
            while self.inflight[index] > 0:
                self.cond.wait()
            scene, self.scenes[index] = self.scenes[index], None
            self.evicting.discard(index)
            self.resident -= scene.nbytes
            self.share = (self.budget - self.resident) / self.nslots
        if free:
            scene.free()

    def _take_batch(self) -> list:
        """The oldest pending request; when more than batch_depth are pending, also the
        later ones of its scene, in arrival order, up to max_views, max_batch_pixels and
        the slot's share of the budget. With max_wait, the requests at the front that have
        waited longer are shed first (later arrivals have waited less)."""
        if self.max_wait is not None:
            now = self._now()
            k = 0
            while k < len(self.pending) and now - self.pending[k].arrival > self.max_wait:
                k += 1
            if k:
                self.shed.extend(self.pending[:k])
                del self.pending[:k]
                if not self.pending:
                    return []
        head = self.pending[0]
        if not self._live(head.scene) or self.working_set([head]) > self.share:
            self.pending.pop(0)
            self.rejected.append(head)
            return []
        batch = [head]
        if len(self.pending) <= self.batch_depth:
            self.pending.pop(0)
            self.inflight[head.scene] += 1
            return batch
        pixels = head.pixels
        for r in self.pending[1:]:
            if len(batch) >= self.max_views:
                break
            if (r.scene == head.scene and pixels + r.pixels <= self.max_batch_pixels
                    and self.working_set(batch + [r]) <= self.share):
                batch.append(r)
                pixels += r.pixels
        taken = set(id(r) for r in batch)
        self.pending = [r for r in self.pending if id(r) not in taken]
        self.inflight[head.scene] += 1
        return batch

    # ── workers ──────────────────────────────────────────────────────────────
    def _now(self) -> float:
        return time.perf_counter() - self.t0

    def _worker(self, slot: _Slot):
        while True:
            with self.cond:
                while not self.pending and not self.stop:
                    self.cond.wait()
                if not self.pending and self.stop:
                    return
                batch = self._take_batch()
            if not batch:
                continue
            try:
                self._render(slot, batch)
            except Exception as e:  # recorded; the run reports it
                with self.cond:
                    self.errors.append(f"slot {slot.index}: {e}")
                    self.rejected.extend(batch)
            finally:
                with self.cond:  # an eviction waits for the scene's batches in flight
                    self.inflight[batch[0].scene] -= 1
                    self.cond.notify_all()

    def _render(self, slot: _Slot, batch: list):
        scene = self.scenes[batch[0].scene]
        t = self._now()
        total = sum(r.pixels for r in batch)
        if len(batch) == 1:  # a lone request takes the one-view path (no batch set-up)
            g, c = scene.struct(), make_camera_struct(batch[0].cam, scene.degree)
            out = C.c_void_p()
            check(lib().msplat_hip_render(slot.ctx.ptr, C.byref(g), C.byref(c), floats(self.bg), C.byref(out)))
            ptr, offs = out.value, [0]
        else:
            views = BatchViews([r.cam for r in batch], scene.degree, [self.bg])
            ptr, offs = render_batch(slot.ctx, scene, views)
        rgba = slot.rgba_buffer(total)
        check(lib().msplat_hip_pack_rgba8(slot.ctx.ptr, C.c_void_p(ptr), total, C.c_void_p(rgba.ptr)))
        slot.ctx.sync()
        done = self._now()
        m = slot.ctx.frame().num_isects if len(batch) == 1 else slot.ctx.batch_isects()
        if m and self.adapt:
            self.ipp[batch[0].scene] = 0.8 * self.ipp[batch[0].scene] + 0.2 * (m / max(total, 1))
        if self.keep_images:
            packed = slot.ctx.download(rgba.ptr, (total, 4), np.uint8)
            for r, o in zip(batch, offs):
                r.image = packed[o:o + r.pixels].reshape(r.cam["H"], r.cam["W"], 4).copy()
        with self.cond:
            for r in batch:
                r.launched, r.done, r.batch_size, r.slot = t, done, len(batch), slot.index
            self.finished.extend(batch)
            self.batches += 1

    # ── a run over an arrival schedule ───────────────────────────────────────
    def run(self, requests: list) -> list:
        """Feeds the requests at their arrival times (seconds from now) and returns them
        completed (or in self.rejected)."""
        threads = [threading.Thread(target=self._worker, args=(s,), daemon=True) for s in self.slots]
        self.finished, self.rejected, self.shed, self.errors, self.batches = [], [], [], [], 0
        self.stop = False
        self.t0 = time.perf_counter()
        for th in threads:
            th.start()
        for r in sorted(requests, key=lambda r: r.arrival):
            delay = r.arrival - self._now()
            if delay > 0:
                time.sleep(delay)
            with self.cond:
                self.pending.append(r)
                self.cond.notify()
        with self.cond:
            self.stop = True
            self.cond.notify_all()
        for th in threads:
            th.join()
        if self.errors:
            raise RuntimeError("; ".join(self.errors))
        return self.finished

    def close(self):
        for s in self.slots:
            s.close()


def serve_plain(scenes: list, requests: list, background=(0.0, 0.0, 0.0), device: int = 0,
                keep_images: bool = False, depth_chunks: int = 0) -> list:
    """The plain form: one msplat_hip_render (+ RGBA8 pack) per request, first come first
    served, on one context; depth_chunks as MSPLAT_TUNE_CHUNKS (0: chosen per frame, 1:
    per tile, as the server renders)."""
    ctx = Context(device)
    ctx.set_tuning(TUNE_CHUNKS, depth_chunks)
    bg = floats(background)
    rgba = None
    t0 = time.perf_counter()
    try:
        for r in sorted(requests, key=lambda r: r.arrival):
            delay = r.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            scene = scenes[r.scene]
            r.launched = time.perf_counter() - t0
            g, c = scene.struct(), make_camera_struct(r.cam, scene.degree)
            out = C.c_void_p()
            check(lib().msplat_hip_render(ctx.ptr, C.byref(g), C.byref(c), bg, C.byref(out)))
            if rgba is None or rgba.nbytes < 4 * r.pixels:
                old, rgba = rgba, DeviceArray(ctx, (int(r.pixels * 1.25) + 1024, 4), np.uint8)
                if old is not None:  # the previous request's pack completed at its sync
                    old.free()
            check(lib().msplat_hip_pack_rgba8(ctx.ptr, out, r.pixels, C.c_void_p(rgba.ptr)))
            ctx.sync()
            r.done = time.perf_counter() - t0
            r.batch_size, r.slot = 1, 0
            if keep_images:
                r.image = ctx.download(rgba.ptr, (r.pixels, 4), np.uint8).reshape(r.cam["H"], r.cam["W"], 4)
        return requests
    finally:
        ctx.sync()
        if rgba is not None:
            rgba.free()
        ctx.close()


def latency_stats(reqs: list, t_first: float | None = None) -> dict:
    """views/s over the run and p50/p99 latency (ms), overall and per resolution."""
    done = [r for r in reqs if r.done == r.done]
    if not done:
        return {"completed": 0}
    lat = np.array([r.done - r.arrival for r in done]) * 1e3
    start = min(r.arrival for r in done) if t_first is None else t_first
    span = max(r.done for r in done) - start
    out = {"completed": len(done), "views_per_s": round(len(done) / max(span, 1e-9), 2),
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "mean_batch": round(float(np.mean([r.batch_size for r in done])), 2)}
    by_res = {}
    for r, l in zip(done, lat):
        by_res.setdefault(f"{r.cam['W']}x{r.cam['H']}", []).append(l)
    out["by_resolution"] = {k: {"n": len(v), "p50_ms": round(float(np.percentile(v, 50)), 3),
                                "p99_ms": round(float(np.percentile(v, 99)), 3)} for k, v in sorted(by_res.items())}
    return out


RESOLUTIONS = [((640, 360), 0.4), ((1280, 720), 0.3), ((1920, 1080), 0.2), ((3840, 2160), 0.1)]


def poisson_requests(rate: float, duration: float, centers: list, seed: int = 0, radius=(3.0, 10.0),
                     resolutions=RESOLUTIONS, focal_scale: float = 0.9) -> list:
    """Poisson arrivals at `rate` per second for `duration` seconds; each request picks a
    scene uniformly, a pose on a sphere of radius U(radius) around the scene's center
    looking at it, and a resolution from the weighted mix (fx = fy = focal_scale W)."""
    from .camera import look_at, make_camera
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    sizes, weights = zip(*resolutions)
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        s = int(rng.integers(0, len(centers)))
        d = rng.normal(0, 1, 3)
        d /= np.linalg.norm(d)
        c = np.asarray(centers[s], np.float64)
        eye = c + d * rng.uniform(*radius)
        W, H = sizes[int(rng.choice(len(sizes), p=np.asarray(weights) / sum(weights)))]
        up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.95 else (1.0, 0.0, 0.0)
        cam = make_camera(look_at(eye, c, up), W, H, focal_scale * W, focal_scale * W, dtu=3)
        out.append(Request(len(out), s, cam, t))
    return out
