"""ctypes bindings to the oracle libraries (test infrastructure).

liboracle.so      oracle/_build — the CPU restatement of the reference path.
libtpt_ref.so     oracle/_ref   — the reference's own sources + harness; only
                  present where the reference tree was available at build time.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtpt_ref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")

RAYS_PRIMARY, RAYS_SECONDARY, RAYS_UNIFORM, RAYS_ADVERSARIAL = 0, 1, 2, 3


def _nthreads():
    return max(1, min(16, os.cpu_count() or 1))


def _perlin_eval(fn, grad, perm, pts, scale, octaves):
    grad = np.ascontiguousarray(grad, np.float32)
    perm = np.ascontiguousarray(perm, np.int32)
    pts = np.ascontiguousarray(pts, np.float32)
    assert grad.size == 768 and perm.size == 768 and pts.ndim == 2 and pts.shape[1] == 3
    n = pts.shape[0]
    noise, turb, tex = (np.zeros(n, np.float32) for _ in range(3))
    args = [grad.ctypes.data, perm.ctypes.data, pts.ctypes.data, C.c_uint64(n), C.c_float(scale)]
    if octaves is not None:
        args.append(C.c_int(octaves))
    fn(*args, noise.ctypes.data, turb.ctypes.data, tex.ctypes.data)
    return noise, turb, tex


def _light_pdf(fn, kind, params, o, d):
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros(o.shape[0], np.float32)
    assert fn(kind, np.asarray(params, np.float32), o.reshape(-1), d.reshape(-1), o.shape[0], out) == 0
    return out


def _light_random(fn, kind, params, o, rtl=None):
    o = np.ascontiguousarray(o, np.float32)
    out = np.zeros(3 * o.shape[0], np.float32)
    args = [kind, np.asarray(params, np.float32), o.reshape(-1), o.shape[0]]
    if rtl is not None:
        args.append(rtl)
    assert fn(*args, out) == 0
    return out.reshape(-1, 3)


def _texture_value(fn, kind, colors, texels, q):
    q = np.ascontiguousarray(q, np.float32)
    out = np.zeros(3 * q.shape[0], np.float32)
    c = np.zeros(6, np.float32) if colors is None else np.asarray(colors, np.float32)
    if texels is None:
        ptr, w, h = None, 0, 0
    else:
        texels = np.ascontiguousarray(texels, np.uint8)
        ptr, w, h = texels.ctypes.data, texels.shape[1], texels.shape[0]
    assert fn(kind, c, ptr, w, h, q.reshape(-1), q.shape[0], out) == 0
    return out.reshape(-1, 3)


def _metal_scatter(fn, albedo, fuzz, q, rtl=None):
    q = np.ascontiguousarray(q, np.float32)
    n = q.shape[0]
    d = np.zeros(3 * n, np.float32)
    ok = np.zeros(n, np.int32)
    args = [np.asarray(albedo, np.float32), fuzz, q.reshape(-1), n]
    if rtl is not None:
        args.append(rtl)
    fn(*args, d, ok)
    return d.reshape(n, 3), ok


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        self.L = L
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_perlin_tables.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_perlin_eval.argtypes = [C.c_void_p] * 3 + [C.c_uint64, C.c_float, C.c_int] + [C.c_void_p] * 3
        L.oracle_scene_build.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(C.c_void_p)]
        L.oracle_scene_destroy.argtypes = [C.c_void_p]
        L.oracle_leaf_count.argtypes = [C.c_void_p]
        L.oracle_leaf_count.restype = C.c_uint32
        L.oracle_scene_settings.argtypes = [C.c_void_p] + [C.POINTER(C.c_uint32)] * 4 + [C.POINTER(C.c_float)]
        L.oracle_camera_fields.argtypes = [C.c_void_p, _f32p]
        L.oracle_trace.argtypes = [C.c_void_p, _f32p, C.c_uint64, C.c_float, C.c_float, _i32p, _f32p,
                                   C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.oracle_make_rays.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, _f32p, C.c_int]
        L.oracle_render.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                    C.c_uint32, C.c_float, C.c_int, C.c_int, _f32p, C.c_void_p, C.c_int]
        L.oracle_philox.argtypes = [C.c_uint32] * 6 + [np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")]
        L.oracle_slab.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float]
        L.oracle_prim_hit.argtypes = [C.c_int, _f32p, _f32p, _f32p, C.c_float, C.c_float, _f32p, _f32p]
        L.oracle_camera_from_params.argtypes = [_f32p, _f32p]
        L.oracle_bvh_dump.argtypes = [C.c_void_p, _f32p, C.c_int]
        L.oracle_sampler.argtypes = [C.c_int, C.c_uint64, C.c_int, _f32p]
        L.oracle_light_pdf.argtypes = [C.c_int, _f32p, _f32p, _f32p, C.c_uint64, _f32p]
        L.oracle_light_random.argtypes = [C.c_int, _f32p, _f32p, C.c_uint64, C.c_int, _f32p]
        L.oracle_texture_value.argtypes = [C.c_int, _f32p, C.c_void_p, C.c_int, C.c_int, _f32p, C.c_uint64, _f32p]
        L.oracle_metal_scatter.argtypes = [_f32p, C.c_float, _f32p, C.c_uint64, C.c_int, _f32p, _i32p]
        L.oracle_pixel_samples.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_uint32, _u32p, C.c_uint64,
                                           C.c_uint32, C.c_uint32, C.c_int, C.c_int, _f32p]
        L.oracle_color_rays.argtypes = [C.c_void_p, _f32p, C.c_uint64, C.c_uint32, C.c_float, C.c_uint64,
                                        C.c_uint32, C.c_int, _f32p]
        L.oracle_guide_forward.argtypes = [_f32p, _f32p, C.c_uint64, _f32p]
        L.oracle_guide_sample.argtypes = [_f32p, _f32p, _f32p, C.c_uint64, _f32p, _f32p]
        L.oracle_guide_pdf.argtypes = [_f32p, _f32p, _f32p, C.c_uint64, _f32p]
        L.oracle_render_guided.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                           C.c_uint32, _f32p, C.c_int, C.c_int, _f32p, C.c_void_p]

    def perlin_tables(self):
        """The scene API's current Perlin tables: grad (256, 3), perm (3, 256)."""
        grad, perm = np.zeros(768, np.float32), np.zeros(768, np.int32)
        self.L.oracle_perlin_tables(grad.ctypes.data, perm.ctypes.data)
        return grad.reshape(256, 3), perm.reshape(3, 256)

    def perlin_eval(self, grad, perm, pts, scale, octaves=5):
        """noise(p), turb(p, octaves), perlin_noise_texture(scale).value(0, 0, p).x() with the given tables."""
        return _perlin_eval(self.L.oracle_perlin_eval, grad, perm, pts, scale, octaves)

    def philox(self, ctr, key):
        out = np.zeros(4, np.uint32)
        self.L.oracle_philox(*[int(c) for c in ctr], *[int(k) for k in key], out)
        return out

    def slab(self, lo, hi, o, d, tmin=0.001, tmax=3.4028234663852886e38):
        f = lambda v: np.asarray(v, np.float32)
        return bool(self.L.oracle_slab(f(lo), f(hi), f(o), f(d), tmin, tmax))

    def prim_hit(self, kind, params, o, d, tmin=0.001, tmax=3.4028234663852886e38):
        """kind: 'sphere' (cx,cy,cz,r) or 'xy'/'xz'/'yz' (a0,a1,b0,b1,k).  Returns (t, rec[8]) or None."""
        k = {"sphere": 0, "xy": 1, "xz": 2, "yz": 3}[kind]
        t = np.zeros(1, np.float32)
        rec = np.zeros(8, np.float32)
        f = lambda v: np.asarray(v, np.float32)
        ok = self.L.oracle_prim_hit(k, f(params), f(o), f(d), tmin, tmax, t, rec)
        return (t[0], rec) if ok == 1 else None

    def camera_from_params(self, params):
        out = np.zeros(24, np.float32)
        self.L.oracle_camera_from_params(np.asarray(params, np.float32), out)
        return out

    def scene(self, name, seed=5489):
        return OracleScene(self, name, seed)

    # -- the guided proposal (no reference counterpart) --
    def guide_forward(self, params, in9):
        in9 = np.ascontiguousarray(in9, np.float32)
        out = np.zeros(40 * in9.shape[0], np.float32)
        self.L.oracle_guide_forward(np.ascontiguousarray(params, np.float32), in9.reshape(-1), in9.shape[0], out)
        return out.reshape(-1, 40)

    def guide_sample(self, params, in9, u):
        """Directions sampled from the decoded mixture of one input with draws
        u (n, 3), and the mixture density at each."""
        u = np.ascontiguousarray(u, np.float32)
        d = np.zeros(3 * u.shape[0], np.float32)
        pdf = np.zeros(u.shape[0], np.float32)
        rc = self.L.oracle_guide_sample(np.ascontiguousarray(params, np.float32), np.asarray(in9, np.float32),
                                        u.reshape(-1), u.shape[0], d, pdf)
        assert rc == 0
        return d.reshape(-1, 3), pdf

    def guide_pdf(self, params, in9, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32)
        pdf = np.zeros(dirs.shape[0], np.float32)
        assert self.L.oracle_guide_pdf(np.ascontiguousarray(params, np.float32), np.asarray(in9, np.float32),
                                       dirs.reshape(-1), dirs.shape[0], pdf) == 0
        return pdf

    # -- single functions (same arguments as Reference's) --
    def sampler(self, kind, n, rtl):
        out = np.zeros(3 * n, np.float32)
        self.L.oracle_sampler(kind, n, rtl, out)
        return out.reshape(n, 3)

    def light_pdf(self, kind, params, o, d):
        return _light_pdf(self.L.oracle_light_pdf, kind, params, o, d)

    def light_random(self, kind, params, o, rtl):
        return _light_random(self.L.oracle_light_random, kind, params, o, rtl)

    def texture_value(self, kind, colors, texels, q):
        return _texture_value(self.L.oracle_texture_value, kind, colors, texels, q)

    def metal_scatter(self, albedo, fuzz, q, rtl):
        return _metal_scatter(self.L.oracle_metal_scatter, albedo, fuzz, q, rtl)


class OracleScene:
    def __init__(self, lib, name, seed):
        self.lib = lib
        self.name = name
        h = C.c_void_p()
        if lib.L.oracle_scene_build(name.encode(), seed, C.byref(h)) != 0:
            raise ValueError(lib.L.oracle_last_error().decode())
        self.h = h
        w, hh, spp, d = (C.c_uint32() for _ in range(4))
        bg = C.c_float()
        lib.L.oracle_scene_settings(h, C.byref(w), C.byref(hh), C.byref(spp), C.byref(d), C.byref(bg))
        self.width, self.height, self.spp, self.max_depth = w.value, hh.value, spp.value, d.value
        self.background = bg.value

    def __del__(self):
        try:
            self.lib.L.oracle_scene_destroy(self.h)
        except Exception:
            pass

    @property
    def n_leaves(self):
        return self.lib.L.oracle_leaf_count(self.h)

    def camera_fields(self):
        out = np.zeros(24, np.float32)
        self.lib.L.oracle_camera_fields(self.h, out)
        return out

    def bvh_dump(self, cap=1 << 20):
        out = np.zeros(6 * cap, np.float32)
        n = self.lib.L.oracle_bvh_dump(self.h, out, cap)
        return out[: 6 * min(n, cap)].reshape(-1, 6)

    def make_rays(self, kind, n, seed=1, nthreads=None):
        rays = np.zeros((6, n), np.float32)
        self.lib.L.oracle_make_rays(self.h, kind, n, seed, rays.reshape(-1), nthreads or _nthreads())
        return rays

    def trace(self, rays, tmin=0.001, tmax=3.4028234663852886e38, record=False, nthreads=None, time=None):
        n = rays.shape[1]
        rays = np.ascontiguousarray(rays, np.float32)
        leaf = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        rec = np.zeros((8, n), np.float32) if record else None
        mat = np.zeros(n, np.int32) if record else None
        time = None if time is None else np.ascontiguousarray(time, np.float32)
        self.lib.L.oracle_trace(self.h, rays.reshape(-1), n, tmin, tmax, leaf, t,
                                rec.ctypes.data if record else None, mat.ctypes.data if record else None,
                                nthreads or _nthreads(), None if time is None else time.ctypes.data)
        return (leaf, t, rec, mat) if record else (leaf, t)

    def pixel_samples(self, cam_params, width, height, pix, spp, max_depth, rtl, nthreads=None):
        """Per-sample de_nan'ed colors (npix, spp, 3) of the listed pixels, each
        pixel from a fresh mt19937 (as Reference.pixel_samples)."""
        pix = np.ascontiguousarray(pix, np.uint32)
        out = np.zeros(pix.shape[0] * spp * 3, np.float32)
        rc = self.lib.L.oracle_pixel_samples(self.h, np.asarray(cam_params, np.float32), width, height,
                                             pix.reshape(-1), pix.shape[0], spp, max_depth, rtl,
                                             nthreads or _nthreads(), out)
        assert rc == 0, "pixel samples need a light list"
        return out.reshape(pix.shape[0], spp, 3)

    def color_rays(self, rays, max_depth=50, background=1.0, seed=1, sample=0, nthreads=None):
        """color() of explicit rays (6, n) with the device's Philox streams; (3, n)."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[1]
        out = np.zeros(3 * n, np.float32)
        self.lib.L.oracle_color_rays(self.h, rays.reshape(-1), n, max_depth, background, seed, sample,
                                     nthreads or _nthreads(), out)
        return out.reshape(3, n)

    def render_guided(self, width, height, spp, params, guide_verts=2, max_depth=None, seed=1, sample_offset=0,
                      nthreads=None):
        """The guided estimator on the device's Philox streams (sums, sumsq)."""
        s = np.zeros(width * height * 3, np.float32)
        sq = np.zeros(width * height * 3, np.float32)
        self.lib.L.oracle_render_guided(self.h, width, height, spp,
                                        max_depth if max_depth is not None else self.max_depth, seed, sample_offset,
                                        np.ascontiguousarray(params, np.float32), guide_verts,
                                        nthreads or _nthreads(), s, sq.ctypes.data)
        return s.reshape(height, width, 3), sq.reshape(height, width, 3)

    def render(self, width, height, spp, max_depth=None, seed=1, sample_offset=0, background=None,
               rng_mode=0, nthreads=None, media_scatter=False):
        s = np.zeros(width * height * 3, np.float32)
        sq = np.zeros(width * height * 3, np.float32)
        self.lib.L.oracle_render(self.h, width, height, spp, max_depth if max_depth is not None else self.max_depth,
                                 seed, sample_offset, self.background if background is None else background,
                                 rng_mode, nthreads or _nthreads(), s, sq.ctypes.data, 1 if media_scatter else 0)
        return s.reshape(height, width, 3), sq.reshape(height, width, 3)


class Reference:
    """The reference's own code (oracle/_ref/libtpt_ref.so)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        L.ref_scene_build.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.ref_perlin_tables.argtypes = [C.c_void_p, C.c_void_p]
        L.ref_perlin_eval.argtypes = [C.c_void_p] * 3 + [C.c_uint64, C.c_float] + [C.c_void_p] * 3
        L.ref_leaf_count.argtypes = [C.c_void_p]
        L.ref_trace.argtypes = [C.c_void_p, _f32p, C.c_uint64, C.c_float, C.c_float, _i32p, _f32p,
                                C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.ref_camera_fields.argtypes = [_f32p, _f32p]
        L.ref_bvh_dump.argtypes = [C.c_void_p, _f32p, C.c_int]
        L.ref_render.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _f32p,
                                 C.c_void_p, C.c_int]
        L.ref_sampler.argtypes = [C.c_int, C.c_uint64, _f32p]
        L.ref_light_pdf.argtypes = [C.c_int, _f32p, _f32p, _f32p, C.c_uint64, _f32p]
        L.ref_light_random.argtypes = [C.c_int, _f32p, _f32p, C.c_uint64, _f32p]
        L.ref_texture_value.argtypes = [C.c_int, _f32p, C.c_void_p, C.c_int, C.c_int, _f32p, C.c_uint64, _f32p]
        L.ref_metal_scatter.argtypes = [_f32p, C.c_float, _f32p, C.c_uint64, _f32p, _i32p]
        L.ref_pixel_samples.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_uint32, _u32p, C.c_uint64, C.c_uint32,
                                        C.c_uint32, C.c_int, _f32p]
        L.ref_perlin_install.argtypes = [C.c_void_p, C.c_void_p]
        L.ref_render_snapshots.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_uint32, C.c_int, _f32p]

    def scene(self, name):
        h = C.c_void_p()
        if self.L.ref_scene_build(name.encode(), C.byref(h)) != 0:
            raise ValueError(f"reference harness: unknown scene {name}")
        return h

    def leaf_count(self, h):
        return self.L.ref_leaf_count(h)

    def trace(self, h, rays, tmin=0.001, tmax=3.4028234663852886e38, record=False, nthreads=None, time=None):
        n = rays.shape[1]
        rays = np.ascontiguousarray(rays, np.float32)
        leaf = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        rec = np.zeros((8, n), np.float32) if record else None
        mat = np.zeros(n, np.int32) if record else None
        time = None if time is None else np.ascontiguousarray(time, np.float32)
        self.L.ref_trace(h, rays.reshape(-1), n, tmin, tmax, leaf, t, rec.ctypes.data if record else None,
                         mat.ctypes.data if record else None, nthreads or _nthreads(),
                         None if time is None else time.ctypes.data)
        return (leaf, t, rec, mat) if record else (leaf, t)

    def camera_fields(self, params):
        out = np.zeros(24, np.float32)
        self.L.ref_camera_fields(np.asarray(params, np.float32), out)
        return out

    def perlin_tables(self):
        """The reference's current perlin_noise static tables: grad (256, 3), perm (3, 256)."""
        grad, perm = np.zeros(768, np.float32), np.zeros(768, np.int32)
        self.L.ref_perlin_tables(grad.ctypes.data, perm.ctypes.data)
        return grad.reshape(256, 3), perm.reshape(3, 256)

    def perlin_eval(self, grad, perm, pts, scale):
        """The reference's noise(p), turb(p) and perlin_noise_texture(scale).value(0, 0, p).x(), given tables."""
        return _perlin_eval(self.L.ref_perlin_eval, grad, perm, pts, scale, None)

    def bvh_dump(self, h, cap=1 << 20):
        out = np.zeros(6 * cap, np.float32)
        n = self.L.ref_bvh_dump(h, out, cap)
        return out[: 6 * min(n, cap)].reshape(-1, 6)

    def render(self, h, cam_params, width, height, spp, max_depth, decorrelate_rows=False, with_sumsq=False):
        s = np.zeros(width * height * 3, np.float32)
        sq = np.zeros(width * height * 3, np.float32)
        rc = self.L.ref_render(h, np.asarray(cam_params, np.float32), width, height, spp, max_depth, s,
                               sq.ctypes.data if with_sumsq else None, 1 if decorrelate_rows else 0)
        if rc != 0:
            raise ValueError("reference render needs a light list")
        if with_sumsq:
            return s.reshape(height, width, 3), sq.reshape(height, width, 3)
        return s.reshape(height, width, 3)

    # -- single functions of the reference (fresh generator where they draw) --
    def sampler(self, kind, n):
        out = np.zeros(3 * n, np.float32)
        self.L.ref_sampler(kind, n, out)
        return out.reshape(n, 3)

    def light_pdf(self, kind, params, o, d):
        return _light_pdf(self.L.ref_light_pdf, kind, params, o, d)

    def light_random(self, kind, params, o):
        return _light_random(self.L.ref_light_random, kind, params, o)

    def texture_value(self, kind, colors, texels, q):
        return _texture_value(self.L.ref_texture_value, kind, colors, texels, q)

    def metal_scatter(self, albedo, fuzz, q):
        return _metal_scatter(self.L.ref_metal_scatter, albedo, fuzz, q)

    def pixel_samples(self, h, cam_params, width, height, pix, spp, max_depth, nthreads=None):
        pix = np.ascontiguousarray(pix, np.uint32)
        out = np.zeros(pix.shape[0] * spp * 3, np.float32)
        rc = self.L.ref_pixel_samples(h, np.asarray(cam_params, np.float32), width, height, pix.reshape(-1),
                                      pix.shape[0], spp, max_depth, nthreads or _nthreads(), out)
        assert rc == 0, "pixel samples need a light list"
        return out.reshape(pix.shape[0], spp, 3)

    def perlin_install(self, grad, perm):
        grad = np.ascontiguousarray(grad, np.float32)
        perm = np.ascontiguousarray(perm, np.int32)
        self.L.ref_perlin_install(grad.ctypes.data, perm.ctypes.data)

    def render_snapshots(self, h, cam_params, width, height, spp, every, max_depth, max_threads=0):
        """The reference's render with its progressive snapshots: (spp / every,
        height, width, 3) running sums after every `every` samples."""
        n = spp // every
        out = np.zeros(n * width * height * 3, np.float32)
        rc = self.L.ref_render_snapshots(h, np.asarray(cam_params, np.float32), width, height, spp, every, max_depth,
                                         max_threads, out)
        if rc != 0:
            raise ValueError("reference snapshots need a light list and spp divisible by every")
        return out.reshape(n, height, width, 3)

