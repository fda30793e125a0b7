"""TEST INFRASTRUCTURE: ctypes front of the f64 CPU oracle (oracle/build/libmb3d_oracle.so).

Only the tests, the smoke check and the benchmark's CPU baseline use it, as
the checker.  The oracle restates the reference's CPU marcher
(src/calc.rs, src/iteration.rs, src/formulas.rs, src/deao.rs); see
mb3d_oracle.cpp for the function-by-function citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libmb3d_oracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "mb3d-rs_amd"))
import mb3d_hip as H  # noqa: E402  (struct layouts of include/mb3d_hip.h)


class OrcCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("pixels", "hits", "steps", "de_evals", "its_p8", "its_box", "its_other")]


class OrcOut(C.Structure):
    _fields_ = [
        ("depth", C.c_void_p),
        ("iters", C.c_void_p),
        ("normals", C.c_void_p),
        ("gbuf", C.c_void_p),
        ("steps", C.c_void_p),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.orc_last_error.restype = C.c_char_p
    lib.orc_params_from_scene.argtypes = [C.POINTER(H.Scene), C.POINTER(H.Params)]
    lib.orc_params_from_scene.restype = C.c_int32
    lib.orc_lighting_default.argtypes = [C.POINTER(H.Lighting)]
    lib.orc_lighting_default.restype = None
    lib.orc_paint_setup.argtypes = [C.POINTER(H.Scene), C.POINTER(H.Params), C.POINTER(H.Lighting),
                                    C.POINTER(H.PaintCamera), C.POINTER(H.LightVals)]
    lib.orc_paint_setup.restype = C.c_int32
    lib.orc_march_rays.argtypes = [C.POINTER(H.Params), C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p,
                                   C.c_int32]
    lib.orc_march_rays.restype = C.c_int32
    lib.orc_de_many.argtypes = [C.POINTER(H.Params), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32]
    lib.orc_de_many.restype = C.c_int32
    lib.orc_mc_batch.argtypes = [C.POINTER(H.Params), C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int32]
    lib.orc_mc_batch.restype = C.c_int32
    lib.orc_gbuf_pack.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    lib.orc_gbuf_pack.restype = None
    lib.orc_gbuf_unpack.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    lib.orc_gbuf_unpack.restype = None
    lib.orc_cos_tab_val.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_int32]
    lib.orc_cos_tab_val.restype = C.c_float
    lib.orc_cos_tables.argtypes = [C.c_void_p]
    lib.orc_cos_tables.restype = None
    lib.orc_paint.argtypes = [C.POINTER(H.PaintCamera), C.POINTER(H.LightVals), C.c_void_p, C.c_int32, C.c_int32,
                              C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32]
    lib.orc_paint.restype = C.c_int32
    lib.orc_pixel_seed.argtypes = [C.c_int64, C.c_int64]
    lib.orc_pixel_seed.restype = C.c_int32
    lib.orc_iterate.argtypes = [C.POINTER(H.Formula), C.c_void_p, C.c_void_p, C.c_double, C.c_int32]
    lib.orc_iterate.restype = None
    lib.orc_de_at.argtypes = [C.POINTER(H.Params), C.c_void_p, C.POINTER(C.c_int32)]
    lib.orc_de_at.restype = C.c_double
    lib.orc_pixel_ray.argtypes = [C.POINTER(H.Params), C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    lib.orc_pixel_ray.restype = None
    lib.orc_render.argtypes = [C.POINTER(H.Params), C.POINTER(H.PostJob), C.c_int32, C.c_int32,
                               C.POINTER(OrcOut), C.c_int32, C.POINTER(OrcCounters)]
    lib.orc_render.restype = C.c_int32
    lib.orc_post.argtypes = [C.POINTER(H.Params), C.POINTER(H.PostJob), C.c_int32, C.c_int32,
                             C.c_void_p, C.c_int32]
    lib.orc_post.restype = C.c_int32
    lib.orc_march_ray.argtypes = [C.POINTER(H.Params), C.c_void_p, C.c_void_p, C.c_double,
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.orc_march_ray.restype = C.c_double
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def params(scene) -> "H.Params":
    p = H.Params()
    r = load().orc_params_from_scene(C.byref(scene), C.byref(p))
    if r != 0:
        raise RuntimeError("oracle params: " + load().orc_last_error().decode())
    return p


def render(p, post=None, row0=0, rows=None, threads=0, cone=None):
    """The oracle's G-buffer of rows [row0, row0+rows): dict of arrays.
    `cone` = (block width, block height, mode): the march with the cone grid
    (mode 1: blocks whose cone passed the end are background; 2: also every
    pixel's march from its block's depth); no post job then."""
    rows = p.height - row0 if rows is None else rows
    n = rows * p.width
    out = {
        "depth": np.zeros(n, np.float64),
        "iters": np.zeros(n, np.uint16),
        "normals": np.zeros((n, 3), np.float64),
        "gbuf": np.zeros((n, 5), np.uint32),
        "steps": np.zeros(n, np.int32),
    }
    o = OrcOut(*(_ptr(out[k]) for k in ("depth", "iters", "normals", "gbuf", "steps")))
    cnt = OrcCounters()
    if cone is not None:
        assert post is None
        r = load().orc_render_cone(C.byref(p), cone[0], cone[1], cone[2], row0, rows, C.byref(o), threads,
                                   C.byref(cnt))
    else:
        r = load().orc_render(C.byref(p), C.byref(post) if post is not None else None, row0, rows,
                              C.byref(o), threads, C.byref(cnt))
    if r != 0:
        raise RuntimeError("oracle render: " + load().orc_last_error().decode())
    out["counters"] = {k: getattr(cnt, k) for k, _ in OrcCounters._fields_}
    return out


def cone_grid(p, bw, bh, threads=0):
    """The cone march's blocks (bw x bh pixels, one level): safe depth per
    block (+inf: background) and the dynamic fog's count, as (rows, cols)."""
    cols, rows = (p.width + bw - 1) // bw, (p.height + bh - 1) // bh
    t = np.zeros(cols * rows)
    cnt = np.zeros(cols * rows, np.float32)
    r = load().orc_cone_grid(C.byref(p), bw, bh, _ptr(t), _ptr(cnt), threads)
    if r != 0:
        raise RuntimeError("oracle cone grid: " + load().orc_last_error().decode())
    return t.reshape(rows, cols), cnt.reshape(rows, cols)


def post(p, job, gbuf, row0=0, rows=None, threads=0):
    rows = p.height - row0 if rows is None else rows
    g = np.ascontiguousarray(gbuf, dtype=np.uint32).copy()
    r = load().orc_post(C.byref(p), C.byref(job), row0, rows, _ptr(g), threads)
    if r != 0:
        raise RuntimeError("oracle post: " + load().orc_last_error().decode())
    return g


def de_at(p, pos):
    a = np.asarray(pos, np.float64)
    its = C.c_int32()
    d = load().orc_de_at(C.byref(p), _ptr(a), C.byref(its))
    return d, its.value


def de_many(p, pos, threads=0):
    """The DE in steps and the iteration count at each of the points (n x 3)."""
    a = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    de = np.zeros(a.shape[0])
    its = np.zeros(a.shape[0], np.int32)
    st = load().orc_de_many(C.byref(p), _ptr(a), a.shape[0], _ptr(de), _ptr(its), threads)
    if st != 0:
        raise RuntimeError(f"orc_de_many failed ({st})")
    return de, its


def deao_probe(p, job, x, y, rec, rnd=None):
    """One pixel's DE AO on its card record (5 words) with what it computed:
    (ray count, amb word, dict of the surface point, each ray's world step,
    each ray's openness times de_mul, each ray's first step factor, and
    step_ao, max_dist, ms_ao, md_d10, de_mul, dsm).  `rnd`: the row
    generator's state before the pixel (the random first step); the state
    after it is in info["rnd"]."""
    out = np.zeros(174)
    amb = C.c_int32()
    r = np.ascontiguousarray(rec, np.uint32)
    st = C.c_int32(0 if rnd is None else int(np.int32(np.uint32(rnd & 0xFFFFFFFF))))
    n = load().orc_deao_probe(C.byref(p), C.byref(job), x, y, _ptr(r), _ptr(out), C.byref(amb),
                              None if rnd is None else C.byref(st))
    if n < 0:
        raise RuntimeError("orc_deao_probe failed")
    info = dict(zip(("step_ao", "max_dist", "ms_ao", "md_d10", "de_mul", "dsm"), out[135:141]))
    info.update(point=out[:3], dirs=out[3:102].reshape(33, 3)[:n], min_ra=out[102:135][:n],
                first=out[141:174][:n], rnd=st.value)
    return n, amb.value, info


def deao_walk(p, job, x, y, rec, ray, rnd=None, cap=4096):
    """The walk of one DE AO ray of pixel (x, y) on its card record: (t, de,
    ms_walk, openness) -- the samples' distances in steps, the DE the walk
    took at each, the marcher's DE stop during the walk and the ray's
    openness (min_ra / de_mul).  `rnd` as for deao_probe."""
    t = np.zeros(cap)
    de = np.zeros(cap)
    out2 = np.zeros(2)
    r = np.ascontiguousarray(rec, np.uint32)
    st = C.c_int32(0 if rnd is None else int(np.int32(np.uint32(rnd & 0xFFFFFFFF))))
    n = load().orc_deao_walk(C.byref(p), C.byref(job), x, y, _ptr(r), None if rnd is None else C.byref(st),
                             ray, cap, _ptr(t), _ptr(de), _ptr(out2))
    if n < 0:
        raise RuntimeError("orc_deao_walk failed")
    return t[:n].copy(), de[:n].copy(), float(out2[0]), float(out2[1])


def color_probe(p, x, y):
    """Pixel (x, y)'s march with what do_color saw at its surface: dict of
    the orbit's end v, the surface point c, rout, rold and the otrap and
    gradient words; None when the pixel is not a surface."""
    out = np.zeros(8)
    words = np.zeros(2, np.uint16)
    r = load().orc_color_probe(C.byref(p), x, y, _ptr(out), _ptr(words))
    if r < 0:
        raise RuntimeError("orc_color_probe: " + load().orc_last_error().decode())
    if r == 0:
        return None
    return {"v": out[:3].copy(), "c": out[3:6].copy(), "rout": out[6], "rold": out[7], "otrap": int(words[0]),
            "gradient": int(words[1])}


def deao_row_seed(y):
    """The DE AO generator's seed of picture row y."""
    return load().orc_deao_row_seed(y)


def deao_rand(seed, n):
    """n draws of the DE AO's generator from `seed`: (states, numbers)."""
    states = np.zeros(n, np.int32)
    vals = np.zeros(n)
    load().orc_deao_rand(int(np.int32(np.uint32(seed & 0xFFFFFFFF))), n, _ptr(states), _ptr(vals))
    return states, vals


def xy_vecs(n):
    """create_xy_vecs_from_normals: two vectors across the normal n."""
    a = np.ascontiguousarray(n, np.float64)
    vx, vy = np.zeros(3), np.zeros(3)
    load().orc_xy_vecs(_ptr(a), _ptr(vx), _ptr(vy))
    return vx, vy


def pixel_ray(p, x, y):
    c = np.zeros(3)
    v = np.zeros(3)
    load().orc_pixel_ray(C.byref(p), x, y, _ptr(c), _ptr(v))
    return c, v


def march_ray(p, origin, direction, zz=0.0):
    o = np.asarray(origin, np.float64)
    d = np.asarray(direction, np.float64)
    its, steps = C.c_int32(), C.c_int32()
    t = load().orc_march_ray(C.byref(p), _ptr(o), _ptr(d), zz, C.byref(its), C.byref(steps))
    return t, its.value, steps.value


def iterate(f, v, j, ade=False, rout=0.0):
    """One Formula::iterate step; `rout` is the stored squared radius that
    Real Power and Bulbox read.  Returns the new (v, j)."""
    v = np.array(v, np.float64)
    j = np.array(j, np.float64)
    load().orc_iterate(C.byref(f), _ptr(v), _ptr(j), float(rout), int(ade))
    return v, j


def pixel_seed(y, x):
    return load().orc_pixel_seed(y, x)


# ---------------------------------------------------------------------------
# the painter


def lighting_default():
    """Lighting::default (src/lighting.rs:290-392), the oracle's restatement."""
    lt = H.Lighting()
    load().orc_lighting_default(C.byref(lt))
    return lt


def paint_setup(s, p, lt=None):
    """(PaintCamera, LightVals): paint_camera + light_vals restated."""
    lt = lighting_default() if lt is None else lt
    cam, lv = H.PaintCamera(), H.LightVals()
    st = load().orc_paint_setup(C.byref(s), C.byref(p), C.byref(lt), C.byref(cam), C.byref(lv))
    if st != 0:
        raise RuntimeError(f"orc_paint_setup failed ({st})")
    return cam, lv


def cos_tab_val(tnr, dotp, rough=0.0, sqr=False):
    return float(load().orc_cos_tab_val(tnr, dotp, rough, int(sqr)))


def cos_tables():
    t = np.zeros((8, 128), np.float32)
    load().orc_cos_tables(t.ctypes.data_as(C.c_void_p))
    return t


def paint(cam, lv, gbuf, x0=0, y0=0, w=None, h=None, fmt=0, threads=0):
    """pixel_color_s over the rectangle: an (h, w, 3) array."""
    w = cam.width if w is None else w
    h = cam.height if h is None else h
    g = np.ascontiguousarray(gbuf, dtype=np.uint32).reshape(-1, 5)
    assert g.shape[0] == w * h
    out = np.zeros((h, w, 3), {0: np.uint8, 1: np.uint16, 2: np.float32}[fmt])
    load().orc_paint(C.byref(cam), C.byref(lv), g.ctypes.data_as(C.c_void_p), x0, y0, w, h, fmt,
                     out.ctypes.data_as(C.c_void_p), threads)
    return out


def gbuf_pack(gbuf):
    """gbuffer::pack of card records: (n, 5) u32 -> n * 18 bytes."""
    g = np.ascontiguousarray(gbuf, dtype=np.uint32).reshape(-1, 5)
    out = np.zeros(g.shape[0] * 18, np.uint8)
    load().orc_gbuf_pack(g.ctypes.data_as(C.c_void_p), g.shape[0], out.ctypes.data_as(C.c_void_p))
    return out


def gbuf_unpack(records):
    """gbuffer::unpack into card records: n * 18 bytes -> (n, 5) u32."""
    r = np.ascontiguousarray(records, dtype=np.uint8).reshape(-1)
    n = r.size // 18
    out = np.zeros((n, 5), np.uint32)
    load().orc_gbuf_unpack(r.ctypes.data_as(C.c_void_p), n, out.ctypes.data_as(C.c_void_p))
    return out


def march_rays(p, tasks, depth_sqrt=False, opts=None, threads=0):
    """SrTracer::march_ray for each task (H.RAY_TASK_DTYPE records), with the
    hits' shadows and occlusion of `opts` (an H.RayOpts)."""
    t = np.ascontiguousarray(tasks, dtype=H.RAY_TASK_DTYPE)
    h = np.zeros(t.shape[0], H.RAY_HIT_DTYPE)
    st = load().orc_march_rays(C.byref(p), t.ctypes.data_as(C.c_void_p), t.shape[0], int(depth_sqrt),
                               C.byref(opts) if opts is not None else None, h.ctypes.data_as(C.c_void_p), threads)
    if st != 0:
        raise RuntimeError(f"orc_march_rays failed ({st})")
    return h


def mc_batch(p, tasks, secant=False, threads=0):
    """McMarch::run for each task (H.MC_TASK_DTYPE records): H.MC_RES_DTYPE
    results."""
    t = np.ascontiguousarray(tasks, dtype=H.MC_TASK_DTYPE)
    r = np.zeros(t.shape[0], H.MC_RES_DTYPE)
    st = load().orc_mc_batch(C.byref(p), t.ctypes.data_as(C.c_void_p), t.shape[0], int(secant),
                             r.ctypes.data_as(C.c_void_p), threads)
    if st != 0:
        raise RuntimeError(f"orc_mc_batch failed ({st}): {load().orc_last_error().decode()}")
    return r
