"""tiny-path-tracer_amd — MI355X wavefront path tracer, Python host side.

Thin ctypes layer over libtpt.so (the C ABI declared in include/tpt.h).  The
scene graph is built by the library's C++ scene API (the reference's
hitable / material / texture / camera classes); Python selects a named scene,
uploads it and drives the two device entry points that replace the
reference's per-ray interfaces:

    Scene.trace(rays)   -> tpt_trace_closest   (hitable::hit, headers/hitable.h:32-33)
    Scene.render(...)   -> tpt_render          (color() + main.cpp's sample loop)

Device buffers are torch tensors on cuda:0 (torch is plumbing only: memory and
streams).  There is no CPU fallback: without a gfx950 device or without the
built library every device call raises.

Import as ``importlib.import_module("tiny-path-tracer_amd")``.
"""
import ctypes as C
import os

import numpy as np

__all__ = ["lib", "HostScene", "Scene", "Guide", "TptError", "scene_names", "build", "tonemap_ppm", "autotune_plan"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtpt.so")

TPT_OK = 0


class TptError(RuntimeError):
    pass


class _Slot(C.Structure):
    _fields_ = [("f", C.c_float * 6), ("a", C.c_uint32), ("b", C.c_uint32)]


class FlatScene(C.Structure):
    _fields_ = [
        ("slots", C.c_void_p), ("n_slots", C.c_uint32),
        ("segments", C.c_void_p), ("n_segments", C.c_uint32),
        ("entries", C.c_void_p), ("n_entries", C.c_uint32),
        ("xforms", C.c_void_p), ("n_xforms", C.c_uint32),
        ("materials", C.c_void_p), ("n_materials", C.c_uint32),
        ("textures", C.c_void_p), ("n_textures", C.c_uint32),
        ("texels", C.c_void_p), ("n_texels", C.c_uint64),
        ("perlin", C.c_void_p),
        ("lights", C.c_void_p), ("n_lights", C.c_uint32), ("light_mode", C.c_uint32),
        ("n_leaves", C.c_uint32), ("max_depth", C.c_uint32), ("bounds", C.c_float * 6),
        ("motions", C.c_void_p), ("n_motions", C.c_uint32), ("segment_xforms", C.c_void_p),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3), ("lower_left", C.c_float * 3), ("horizontal", C.c_float * 3),
        ("vertical", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3), ("w", C.c_float * 3),
        ("lens_radius", C.c_float), ("time0", C.c_float), ("time1", C.c_float),
    ]

    def as_array(self):
        vals = []
        for name in ("origin", "lower_left", "horizontal", "vertical", "u", "v", "w"):
            vals.extend(getattr(self, name))
        vals += [self.lens_radius, self.time0, self.time1]
        return np.array(vals, np.float32)


class SceneInfo(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("background", C.c_float), ("n_leaves", C.c_uint32), ("lookfrom", C.c_float * 3),
                ("lookat", C.c_float * 3), ("vfov", C.c_float), ("aperture", C.c_float), ("focus", C.c_float),
                ("time0", C.c_float), ("time1", C.c_float)]


class RenderDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("max_depth", C.c_uint32),
                ("seed", C.c_uint64), ("sample_offset", C.c_uint32), ("background", C.c_float),
                ("guiding", C.c_uint32), ("queue_capacity", C.c_uint32), ("flags", C.c_uint32),
                ("guide_vertices", C.c_uint32)]


class Job(C.Structure):
    """tpt_job: one render of a pooled scene."""
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("sample_offset", C.c_uint32), ("seed", C.c_uint64),
                ("background", C.c_float), ("reserved", C.c_uint32), ("camera", Camera), ("d_sum", C.c_void_p),
                ("d_sumsq", C.c_void_p)]


class Launch(C.Structure):
    """tpt_launch: settings of one tpt_render_jobs call."""
    _fields_ = [("queue_capacity", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("graph_iters", C.c_uint32),
                ("flags", C.c_uint32), ("guiding", C.c_uint32), ("guide_vertices", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("segments", C.c_uint64), ("light_queries", C.c_uint64),
                ("nan_samples", C.c_uint64), ("kernel_ms", C.c_double), ("nodes_visited", C.c_uint64),
                ("trace_ms", C.c_double), ("shade_ms", C.c_double), ("trace_launches", C.c_uint64),
                ("shade_launches", C.c_uint64), ("tail_segments", C.c_uint64), ("tail_ms", C.c_double),
                ("tail_launches", C.c_uint64), ("box_tests", C.c_uint64), ("sphere_tests", C.c_uint64),
                ("rect_tests", C.c_uint64), ("xforms", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_hip_lib = None


def _hip():
    """The HIP runtime (device-to-device copies of guide rows), loaded once."""
    global _hip_lib
    if _hip_lib is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _hip_lib = h
    return _hip_lib


def lib():
    """Loads libtpt.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TptError(f"{LIB_PATH} not built: run `make -C tiny-path-tracer_amd` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, f32, i32p = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_void_p
    L.tpt_abi_version.restype = C.c_int
    L.tpt_last_error.restype = C.c_char_p
    L.tpt_scene_names.restype = C.c_char_p
    L.tpt_device_count.restype = C.c_int
    L.tpt_host_scene_build.argtypes = [C.c_char_p, u32, C.POINTER(vp)]
    L.tpt_host_scene_flat.argtypes = [vp, C.POINTER(FlatScene)]
    L.tpt_host_scene_camera.argtypes = [vp, C.POINTER(Camera)]
    L.tpt_host_scene_info.argtypes = [vp, C.POINTER(SceneInfo)]
    L.tpt_host_scene_destroy.argtypes = [vp]
    L.tpt_camera_make.argtypes = [vp, vp, vp, f32, f32, f32, f32, f32, f32, C.POINTER(Camera)]
    L.tpt_scene_create.argtypes = [C.POINTER(FlatScene), C.POINTER(vp)]
    L.tpt_scene_destroy.argtypes = [vp]
    L.tpt_trace_closest.argtypes = [vp, vp, u64, f32, f32, i32p, vp, vp]
    L.tpt_trace_closest_record.argtypes = [vp, vp, u64, f32, f32, i32p, vp, vp, i32p, vp]
    L.tpt_trace_closest_timed.argtypes = [vp, vp, vp, u64, f32, f32, i32p, vp, vp, i32p, vp]
    L.tpt_render.argtypes = [vp, C.POINTER(Camera), C.POINTER(RenderDesc), vp, vp, C.POINTER(RenderStats), vp]
    L.tpt_render_guided.argtypes = [vp, vp, C.POINTER(Camera), C.POINTER(RenderDesc), vp, vp,
                                    C.POINTER(RenderStats), vp]
    L.tpt_scene_create_pool.argtypes = [C.POINTER(FlatScene), u32, C.POINTER(vp)]
    L.tpt_render_jobs.argtypes = [vp, C.POINTER(Job), u32, C.POINTER(Launch), vp, C.POINTER(RenderStats), vp]
    L.tpt_autotune.argtypes = [vp, C.POINTER(Job), u32, vp, u32, C.POINTER(Launch), C.POINTER(C.c_float)]
    L.tpt_autotune_plan.argtypes = [C.POINTER(Job), u32, C.POINTER(Launch), u32, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.tpt_guide_create.argtypes = [u64, C.POINTER(vp)]
    L.tpt_guide_destroy.argtypes = [vp]
    L.tpt_guide_get_params.argtypes = [vp, vp]
    L.tpt_guide_set_params.argtypes = [vp, vp]
    L.tpt_guide_infer.argtypes = [vp, vp, u64, vp, vp]
    L.tpt_guide_gradient.argtypes = [vp, vp, u64, vp, C.POINTER(C.c_float), vp]
    L.tpt_guide_train.argtypes = [vp, vp, u64, u64, u32, f32, C.POINTER(C.c_float), vp]
    L.tpt_guide_batch.argtypes = [vp, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
    _lib = L
    return L


def _check(rc, what):
    if rc != TPT_OK:
        raise TptError(f"{what} failed ({rc}): {lib().tpt_last_error().decode()}")


def scene_names():
    return lib().tpt_scene_names().decode().split(",")


class HostScene:
    """A named scene built and flattened on the host (no device needed)."""

    def __init__(self, name, seed=5489):
        L = lib()
        self.name = name
        h = C.c_void_p()
        _check(L.tpt_host_scene_build(name.encode(), seed, C.byref(h)), f"tpt_host_scene_build({name})")
        self.h = h
        self.flat = FlatScene()
        _check(L.tpt_host_scene_flat(h, C.byref(self.flat)), "tpt_host_scene_flat")
        self.camera = Camera()
        _check(L.tpt_host_scene_camera(h, C.byref(self.camera)), "tpt_host_scene_camera")
        self.info = SceneInfo()
        _check(L.tpt_host_scene_info(h, C.byref(self.info)), "tpt_host_scene_info")

    def material_kinds(self):
        """Kind (TPT_MAT_*) of every flattened material, by material index."""
        n = self.flat.n_materials
        if n == 0:
            return np.zeros(0, np.uint32)
        words = np.ctypeslib.as_array(C.cast(self.flat.materials, C.POINTER(C.c_uint32)), shape=(n * 8,))
        return words[0::8].copy()

    def slots(self):
        n = self.flat.n_slots
        arr = (_Slot * n).from_address(self.flat.slots) if n else []
        return [(tuple(s.f), s.a, s.b) for s in arr]

    def __del__(self):
        try:
            lib().tpt_host_scene_destroy(self.h)
        except Exception:
            pass


def _torch():
    import torch

    return torch


def _check_sums(buf, shape, what):
    """A caller-supplied accumulator must be what the kernels index: a
    contiguous float32 device tensor of shape (H, W, 3)."""
    torch = _torch()
    if not isinstance(buf, torch.Tensor):
        raise TptError(f"{what}: expected a torch tensor, got {type(buf).__name__}")
    if buf.dtype != torch.float32:
        raise TptError(f"{what}: dtype must be float32, not {buf.dtype}")
    if tuple(buf.shape) != tuple(shape):
        raise TptError(f"{what}: shape {tuple(buf.shape)} does not match the image {tuple(shape)}")
    if not buf.is_contiguous():
        raise TptError(f"{what}: must be contiguous")
    if not buf.is_cuda:
        raise TptError(f"{what}: must live on the device")


class Scene:
    """A scene (or a pool of scenes) uploaded to the current HIP device."""

    def __init__(self, host_scene, pool=None):
        L = lib()
        if L.tpt_device_count() < 1:
            raise TptError("no gfx950 device visible: the device path has no CPU fallback")
        self.hosts = list(pool) if pool is not None else [host_scene]
        self.host = self.hosts[0]
        h = C.c_void_p()
        if pool is None:
            _check(L.tpt_scene_create(C.byref(host_scene.flat), C.byref(h)), "tpt_scene_create")
        else:
            arr = (FlatScene * len(self.hosts))(*[hs.flat for hs in self.hosts])
            _check(L.tpt_scene_create_pool(arr, len(self.hosts), C.byref(h)), "tpt_scene_create_pool")
        self.h = h

    @classmethod
    def named(cls, name, seed=5489):
        return cls(HostScene(name, seed))

    @classmethod
    def pool(cls, host_scenes):
        """Several HostScenes in one handle (tpt_scene_create_pool)."""
        return cls(None, pool=host_scenes)

    def job(self, scene=0, width=None, height=None, spp=None, max_depth=None, seed=1, sample_offset=0,
            background=None, camera=None, sum_buf=None, sumsq_buf=None):
        """A Job for render_jobs, with zeroed device sums unless given; returns
        (job, sum, sumsq)."""
        torch = _torch()
        info = self.hosts[scene].info
        j = Job()
        j.scene = scene
        j.width = width or info.width
        j.height = height or info.height
        j.spp = info.spp if spp is None else spp
        j.max_depth = info.max_depth if max_depth is None else max_depth
        j.seed = seed
        j.sample_offset = sample_offset
        j.background = info.background if background is None else background
        j.camera = camera if camera is not None else self.hosts[scene].camera
        shape = (j.height, j.width, 3)
        if sum_buf is None:
            sum_buf = torch.zeros(shape, dtype=torch.float32, device="cuda")
        if sumsq_buf is None:
            sumsq_buf = torch.zeros(shape, dtype=torch.float32, device="cuda")
        _check_sums(sum_buf, shape, "sum_buf")
        if sumsq_buf is not False:
            _check_sums(sumsq_buf, shape, "sumsq_buf")
        j.d_sum = sum_buf.data_ptr()
        j.d_sumsq = sumsq_buf.data_ptr() if sumsq_buf is not False else None
        return j, sum_buf, sumsq_buf

    def render_jobs(self, jobs, launch=None, guide=None, stream=None):
        """tpt_render_jobs over a list of Job; returns the stats dict."""
        torch = _torch()
        arr = (Job * len(jobs))(*jobs)
        stats = RenderStats()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(lib().tpt_render_jobs(self.h, arr, len(jobs), C.byref(launch) if launch is not None else None,
                                     guide.h if guide is not None else None, C.byref(stats), s), "tpt_render_jobs")
        return stats.as_dict()

    def autotune(self, jobs, guide=None, calib_spp=2, base=None):
        """tpt_autotune: returns (Launch, calibration ms)."""
        arr = (Job * len(jobs))(*jobs)
        best = base if base is not None else Launch()
        ms = C.c_float()
        _check(lib().tpt_autotune(self.h, arr, len(jobs), guide.h if guide is not None else None, calib_spp,
                                  C.byref(best), C.byref(ms)), "tpt_autotune")
        return best, ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().tpt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace(self, rays, tmin=0.001, tmax=3.4028234663852886e38, record=False, stream=None, time=None):
        """rays: float32 tensor (6, n) on the device (ox, oy, oz, dx, dy, dz);
        time: optional float32 tensor (n) of ray times (moving spheres).
        Returns (leaf int32[n], t float32[n]) [+ (rec float32[8, n], mat int32[n])]."""
        torch = _torch()
        assert rays.is_cuda and rays.dtype == torch.float32 and rays.shape[0] == 6
        rays = rays.contiguous()
        n = rays.shape[1]
        leaf = torch.empty(n, dtype=torch.int32, device=rays.device)
        t = torch.empty(n, dtype=torch.float32, device=rays.device)
        s = stream if stream is not None else torch.cuda.current_stream(rays.device).cuda_stream
        if time is not None:
            time = time.contiguous()
            assert time.is_cuda and time.dtype == torch.float32 and time.shape == (n,)
            rec = torch.empty((8, n), dtype=torch.float32, device=rays.device) if record else None
            mat = torch.empty(n, dtype=torch.int32, device=rays.device) if record else None
            _check(lib().tpt_trace_closest_timed(self.h, rays.data_ptr(), time.data_ptr(), n, tmin, tmax,
                                                 leaf.data_ptr(), t.data_ptr(), rec.data_ptr() if record else None,
                                                 mat.data_ptr() if record else None, s),
                   "tpt_trace_closest_timed")
            return (leaf, t, rec, mat) if record else (leaf, t)
        if record:
            rec = torch.empty((8, n), dtype=torch.float32, device=rays.device)
            mat = torch.empty(n, dtype=torch.int32, device=rays.device)
            _check(lib().tpt_trace_closest_record(self.h, rays.data_ptr(), n, tmin, tmax, leaf.data_ptr(),
                                                  t.data_ptr(), rec.data_ptr(), mat.data_ptr(), s),
                   "tpt_trace_closest_record")
            return leaf, t, rec, mat
        _check(lib().tpt_trace_closest(self.h, rays.data_ptr(), n, tmin, tmax, leaf.data_ptr(), t.data_ptr(), s),
               "tpt_trace_closest")
        return leaf, t

    def render(self, width=None, height=None, spp=None, max_depth=None, seed=1, sample_offset=0, background=None,
               guiding=0, queue_capacity=0, sum_buf=None, sumsq_buf=None, stream=None, camera=None,
               profile=False, guide=None, guide_vertices=0, media_scatter=False, tail=True, count_ops=False):
        """Renders into (and returns) device tensors sum, sumsq of shape (H, W, 3);
        row 0 is the bottom image row as in the reference.  tail=False drains
        every pass with the bounce loop alone (TPT_RENDER_NO_TAIL); count_ops
        counts the walk's operations by kind (TPT_RENDER_COUNT_OPS)."""
        torch = _torch()
        info = self.host.info
        desc = RenderDesc()
        desc.width = width or info.width
        desc.height = height or info.height
        desc.spp = info.spp if spp is None else spp
        desc.max_depth = info.max_depth if max_depth is None else max_depth
        desc.seed = seed
        desc.sample_offset = sample_offset
        desc.background = info.background if background is None else background
        desc.guiding = guiding
        desc.queue_capacity = queue_capacity
        desc.flags = (1 if profile else 0) | (2 if media_scatter else 0) | (0 if tail else 4) | (8 if count_ops else 0)
        desc.guide_vertices = guide_vertices
        shape = (desc.height, desc.width, 3)
        if sum_buf is None:
            sum_buf = torch.zeros(shape, dtype=torch.float32, device="cuda")
        if sumsq_buf is None:
            sumsq_buf = torch.zeros(shape, dtype=torch.float32, device="cuda")
        _check_sums(sum_buf, shape, "sum_buf")
        _check_sums(sumsq_buf, shape, "sumsq_buf")
        stats = RenderStats()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        cam = camera if camera is not None else self.host.camera
        if (width and width != info.width) or (height and height != info.height):
            if camera is None:
                raise TptError("a different resolution needs its own camera (aspect ratio)")
        if guide is not None:
            _check(lib().tpt_render_guided(self.h, guide.h, C.byref(cam), C.byref(desc), sum_buf.data_ptr(),
                                           sumsq_buf.data_ptr(), C.byref(stats), s), "tpt_render_guided")
        else:
            _check(lib().tpt_render(self.h, C.byref(cam), C.byref(desc), sum_buf.data_ptr(), sumsq_buf.data_ptr(),
                                    C.byref(stats), s), "tpt_render")
        return sum_buf, sumsq_buf, stats.as_dict()


NPARAM = 14576
ROW = 16


class Guide:
    """The online-trained guiding network (tpt_guide_* in include/tpt.h)."""

    def __init__(self, seed=1):
        L = lib()
        h = C.c_void_p()
        _check(L.tpt_guide_create(seed, C.byref(h)), "tpt_guide_create")
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().tpt_guide_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def get_params(self):
        out = np.zeros(NPARAM + 1, np.float32)
        _check(lib().tpt_guide_get_params(self.h, out.ctypes.data), "tpt_guide_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (NPARAM + 1,)
        _check(lib().tpt_guide_set_params(self.h, p.ctypes.data), "tpt_guide_set_params")

    def infer(self, inputs):
        """inputs: (n, 9) float32 device tensor -> (n, 40) raw outputs."""
        torch = _torch()
        inputs = inputs.contiguous()
        out = torch.empty((inputs.shape[0], 40), dtype=torch.float32, device=inputs.device)
        _check(lib().tpt_guide_infer(self.h, inputs.data_ptr(), inputs.shape[0], out.data_ptr(),
                                     torch.cuda.current_stream().cuda_stream), "tpt_guide_infer")
        return out

    def gradient(self, rows):
        torch = _torch()
        rows = rows.contiguous()
        grad = torch.empty(NPARAM + 1, dtype=torch.float32, device=rows.device)
        loss = C.c_float()
        _check(lib().tpt_guide_gradient(self.h, rows.data_ptr(), rows.shape[0], grad.data_ptr(), C.byref(loss),
                                        torch.cuda.current_stream().cuda_stream), "tpt_guide_gradient")
        return grad, loss.value

    def train(self, rows=None, n_batch=1 << 18, iters=1, lr=1e-3, wait=True):
        """Adam steps on `rows` ((n, 16) device tensor) or, if None, on the rows
        recorded by the last guided render, on the current stream.  Returns the
        last batch loss; wait=False queues the steps and returns None at once.
        Steps on one guide share its gradient buffers: queue them on one
        stream at a time."""
        torch = _torch()
        loss = C.c_float()
        ptr, n = (rows.data_ptr(), rows.shape[0]) if rows is not None else (None, 0)
        _check(lib().tpt_guide_train(self.h, ptr, n, n_batch, iters, lr, C.byref(loss) if wait else None,
                                     torch.cuda.current_stream().cuda_stream), "tpt_guide_train")
        return loss.value if wait else None

    def copy_rows(self, out=None):
        """The rows recorded by the last guided render, copied on the current
        stream into `out` (reused when large enough) or a new tensor."""
        torch = _torch()
        ptr = C.c_void_p()
        n = C.c_uint64()
        _check(lib().tpt_guide_batch(self.h, C.byref(ptr), C.byref(n)), "tpt_guide_batch")
        if out is None or out.shape[0] < n.value:
            out = torch.empty((n.value, ROW), dtype=torch.float32, device="cuda")
        if n.value:
            rc = _hip().hipMemcpyAsync(C.c_void_p(out.data_ptr()), ptr, n.value * ROW * 4, 3,
                                    C.c_void_p(torch.cuda.current_stream().cuda_stream))
            if rc != 0:
                raise TptError(f"hipMemcpyAsync failed ({rc})")
        return out[:n.value]

    def batch_rows(self):
        """Copy of the rows recorded by the last guided render, (n, 16) float32 on the device."""
        torch = _torch()
        ptr = C.c_void_p()
        n = C.c_uint64()
        _check(lib().tpt_guide_batch(self.h, C.byref(ptr), C.byref(n)), "tpt_guide_batch")
        out = torch.empty((n.value, ROW), dtype=torch.float32, device="cuda")
        if n.value:
            torch.cuda.synchronize()
            rc = _hip().hipMemcpy(C.c_void_p(out.data_ptr()), ptr, n.value * ROW * 4, 3)  # device to device
            if rc != 0:
                raise TptError(f"hipMemcpy failed ({rc})")
        return out


class ProgressiveRender:
    """A frame refined pass by pass, as the reference's progressive "bonus"
    snapshots (main.cpp:111-133): every pass adds samples [spp, spp + n) to
    the same sums, so any prefix of passes is an unbiased estimate.

    With a Guide, passes run the guided estimator; a pass with train=True also
    records training rows from every record_stride-th path and is followed by
    `iters` Adam steps of `batch` rows on them (the online training loop).
    Every sample is kept, including those of the early, barely trained passes:
    each is an unbiased sample of the same integral.

    overlap=True trains on a second stream while the next pass renders: the
    recorded rows are copied there, the Adam steps queued behind the copy, and
    the render continues at once with whatever weights the guide holds when
    each block starts (a block stages them once, so its sampling and its
    densities use the same network, and the estimate stays unbiased).  The next
    recording pass waits for the copy before it overwrites the rows.  Results
    then depend on timing."""

    def __init__(self, scene, guide=None, seed=1, record_stride=8, iters=16, batch=1 << 18, lr=1e-3,
                 guide_vertices=2, width=None, height=None, camera=None, overlap=False):
        torch = _torch()
        info = scene.host.info
        self.scene, self.guide, self.seed = scene, guide, seed
        self.width, self.height = width or info.width, height or info.height
        self.camera = camera
        self.record_stride, self.iters, self.batch, self.lr = record_stride, iters, batch, lr
        self.guide_vertices = guide_vertices
        shape = (self.height, self.width, 3)
        self.sum = torch.zeros(shape, dtype=torch.float32, device="cuda")
        self.sumsq = torch.zeros(shape, dtype=torch.float32, device="cuda")
        self.spp = 0
        self.train_loss = None
        self.overlap = overlap
        if overlap:
            self._stream = torch.cuda.Stream()
            self._copied = None
            self._rows = None

    def add(self, spp, train=False):
        """Renders spp more samples per pixel (train: record rows and train the
        guide afterwards).  Returns the render stats."""
        torch = _torch()
        if train and self.guide is None:
            raise TptError("training needs a Guide")
        if self.overlap and train and self._copied is not None:
            torch.cuda.current_stream().wait_event(self._copied)  # the last rows are out of the batch
        _, _, st = self.scene.render(width=self.width, height=self.height, spp=spp, seed=self.seed,
                                     sample_offset=self.spp, sum_buf=self.sum, sumsq_buf=self.sumsq,
                                     camera=self.camera, guide=self.guide,
                                     guiding=self.record_stride if train else 0,
                                     guide_vertices=self.guide_vertices)
        self.spp += spp
        if train and self.overlap:
            self._stream.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(self._stream):
                self._rows = self.guide.copy_rows(self._rows)
                self._copied = torch.cuda.Event()
                self._copied.record()
                self.guide.train(self._rows, n_batch=self.batch, iters=self.iters, lr=self.lr, wait=False)
        elif train:
            self.train_loss = self.guide.train(None, n_batch=self.batch, iters=self.iters, lr=self.lr)
        return st

    def finish(self):
        """Waits for training queued by overlap mode."""
        if self.overlap:
            self._stream.synchronize()

    def mean(self):
        return self.sum / max(self.spp, 1)


def autotune_plan(jobs, base=None, calib_spp=2):
    """tpt_autotune_plan (host only): the distinct queue capacities the
    autotuner times and the spp cap of its calibration jobs."""
    arr = (Job * len(jobs))(*jobs)
    caps = (C.c_uint32 * 6)()  # TPT_AUTOTUNE_MAX_CAPS
    n, k = C.c_uint32(), C.c_uint32()
    _check(lib().tpt_autotune_plan(arr, len(jobs), C.byref(base) if base is not None else None, calib_spp, caps,
                                   C.byref(n), C.byref(k)), "tpt_autotune_plan")
    return list(caps[:n.value]), k.value


def display_rmse(sum_buf, spp, reference_display):
    """RMSE over all pixels and channels of clamp(sqrt(mean), 0, 1), the
    reference's display transform (main.cpp:135-136), against a display image."""
    torch = _torch()
    d = torch.clamp(torch.sqrt(sum_buf / spp), 0.0, 1.0)
    return torch.sqrt(torch.mean((d - reference_display) ** 2)).item()


def camera_make(lookfrom, lookat, vup=(0.0, 1.0, 0.0), vfov=90.0, aspect=1.0, aperture=0.0, focus=1.0, t0=0.0,
                t1=0.0):
    """tpt_camera_make: the fields of camera_with_blur(...) (src/camera.cc:2-21)."""
    f3 = C.c_float * 3
    cam = Camera()
    _check(lib().tpt_camera_make(f3(*lookfrom), f3(*lookat), f3(*vup), vfov, aspect, aperture, focus, t0, t1,
                                 C.byref(cam)), "tpt_camera_make")
    return cam


def tonemap_ppm(mean_rgb):
    """The reference's output transform (main.cpp:135-142, 176-189): sqrt,
    int(255.99f·c), clamp [0,255], rows written top first.  mean_rgb: float
    array (H, W, 3) with row 0 at the bottom.  Returns the P3 text."""
    c = np.sqrt(mean_rgb.astype(np.float32))
    q = (np.float32(255.99) * c)
    q = np.where(np.isfinite(q), q, np.float32(-1.0))
    iv = np.clip(np.trunc(q), 0, 255).astype(np.int32)
    h, w, _ = iv.shape
    lines = [f"P3\n{w} {h}\n255\n"]
    for j in range(h - 1, -1, -1):
        row = iv[j]
        lines.append("".join(f"{r} {g} {b} \n" for r, g, b in row))
    return "".join(lines)


def build():
    """Compiles libtpt.so in-tree (hipcc for gfx950 + g++)."""
    import subprocess

    subprocess.run(["make", "-C", _HERE, "-j8"], check=True)
