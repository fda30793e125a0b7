"""Parity metrics of a G-buffer against the f64 oracle's (test and bench
infrastructure).  Tolerances (the target's): hit/miss mask agreement >= 99.9 %,
relative depth error <= 1e-4 of the scene scale, normal error < 0.5 degrees,
iteration counts equal away from silhouette pixels (a pixel whose 8
neighbours do not all share its mask)."""
from __future__ import annotations

import numpy as np


def masks(depth):
    d = np.asarray(depth, np.float64)
    obj = np.isfinite(d)          # object (surface or inside at the start)
    hit = obj & (d > 0)           # surface found by the march
    return obj, hit


def silhouette(obj, w, h):
    m = obj.reshape(h, w)
    pad = np.pad(m, 1, mode="edge")
    sil = np.zeros_like(m)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dx or dy:
                sil |= pad[1 + dy:1 + dy + h, 1 + dx:1 + dx + w] != m
    return sil.ravel()


def compare(gpu: dict, ref: dict, w: int, h: int, step_width: float, scale: float) -> dict:
    og, hg = masks(gpu["depth64"] if gpu.get("depth64") is not None else gpu["depth"])
    oo, ho = masks(ref["depth"])
    both = hg & ho
    sil = silhouette(oo, w, h) | silhouette(og, w, h)
    g64 = np.asarray(gpu["depth64"] if gpu.get("depth64") is not None else gpu["depth"], np.float64)
    rel = np.abs(g64[both] - ref["depth"][both]) * step_width / scale
    ng = np.asarray(gpu["normals"], np.float64)[both]
    no = np.asarray(ref["normals"], np.float64)[both]
    lg, lo = np.linalg.norm(ng, axis=1), np.linalg.norm(no, axis=1)
    cosang = np.clip(np.sum(ng * no, axis=1) / np.maximum(lg * lo, 1e-300), -1.0, 1.0)
    ang = np.degrees(np.arccos(cosang))
    # a flat DE (a zero gradient, e.g. inside both parts of a smooth DE
    # combination) gives the zero normal on both sides: that agrees
    ang[(lg == 0) & (lo == 0)] = 0.0
    inner = both & ~sil
    it_eq = (gpu["iters"][inner] == ref["iters"][inner])
    q = lambda a, p: float(np.percentile(a, p)) if a.size else 0.0
    return {
        "pixels": int(og.size),
        "mask_agree": float((og == oo).mean()),
        "hits_gpu": int(hg.sum()),
        "hits_ref": int(ho.sum()),
        "both": int(both.sum()),
        "depth_rel_max": float(rel.max()) if rel.size else 0.0,
        "depth_rel_p99": q(rel, 99),
        "depth_rel_mean": float(rel.mean()) if rel.size else 0.0,
        "normal_deg_max": float(ang.max()) if ang.size else 0.0,
        "normal_deg_p99": q(ang, 99),
        "normal_deg_mean": float(ang.mean()) if ang.size else 0.0,
        "depth_ok_frac": float((rel <= 1e-4).mean()) if rel.size else 1.0,
        "normal_ok_frac": float((ang < 0.5).mean()) if ang.size else 1.0,
        "iters_equal_inner": float(it_eq.mean()) if it_eq.size else 1.0,
        "inner": int(inner.sum()),
    }


def card_metrics(g_dev, g_ref) -> dict:
    """The reference's own card-against-CPU measures (gbuffer_matches_cpu,
    src/gpu.rs:4178-4244) on two G-buffers of 5-word card records: same
    background/object share, colour values (orbit trap, 15 bits) off by more
    than 500, flag bits (bit 15 of the trap and the gradient) that differ,
    mean |zpos difference| (of 32768), mean normal cosine and the share of
    normals off by more than 25 degrees (cosine < 0.9)."""
    a = np.asarray(g_dev, np.uint32).reshape(-1, 5)
    b = np.asarray(g_ref, np.uint32).reshape(-1, 5)
    bg_a, bg_b = (a[:, 2] >> 16) >= 32768, (b[:, 2] >> 16) >= 32768
    both = ~bg_a & ~bg_b
    ot_a, ot_b = (a[both, 3] >> 16), (b[both, 3] >> 16)
    gr_a, gr_b = (a[both, 3] & 0xFFFF), (b[both, 3] & 0xFFFF)
    col_off = np.abs((ot_a & 0x7FFF).astype(np.int64) - (ot_b & 0x7FFF).astype(np.int64)) > 500
    flag_off = (((ot_a ^ ot_b) & 0x8000) != 0) | (((gr_a ^ gr_b) & 0x8000) != 0)
    zd = np.abs((a[both, 2] >> 16).astype(np.int64) - (b[both, 2] >> 16).astype(np.int64))

    def normals(g):
        n0 = (g[:, 0] & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.float64)
        n1 = (g[:, 0] >> 16).astype(np.uint16).view(np.int16).astype(np.float64)
        n2 = (g[:, 1] & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.float64)
        return np.stack([n0, n1, n2], 1) / 32767.0

    dots = np.sum(normals(a[both]) * normals(b[both]), 1)
    nb = max(int(both.sum()), 1)
    return {
        "same": float((bg_a == bg_b).mean()),
        "colour_off": float(col_off.sum()) / nb,
        "flags_off": float(flag_off.sum()) / nb,
        "zpos_diff": float(zd.sum()) / nb,
        "normal_cos": float(dots.sum()) / nb,
        "normals_bad": float((dots < 0.9).sum()) / nb,
    }


def card_bar(m: dict) -> bool:
    """gbuffer_matches_cpu's asserts."""
    return (m["same"] > 0.99 and m["colour_off"] < 0.03 and m["flags_off"] < 0.01 and m["zpos_diff"] < 20.0
            and m["normal_cos"] > 0.97 and m["normals_bad"] < 0.03)
