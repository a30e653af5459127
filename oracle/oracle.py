"""FP64 CPU oracle for the SAR backprojection hot path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker: the product path in
``bp_project_amd`` never calls it and fails loudly without its HIP library.

Layout
------
* ``liborc.so`` (built from ``bp_oracle.c`` by ``oracle/Makefile``) holds the
  FP64 restatements of the reference's Numba kernels, each citing the
  ``fast_bp/ffbp.py`` lines it follows.  This module wraps them with ctypes.
* ``form_ffbp`` restates the recursive driver of ``fast_bp/ffbp.py:114-180``
  in Python (node geometry with the same numpy expressions), calling the C
  kernels for leaves and regrids.  Two storage policies:
    - ``"parity"``: complex64 after every leaf/regrid and a complex64 ``+=``,
      exactly the reference's rounding points (ffbp.py:45,60,81,165-167);
    - ``"fp64"``:   complex128 end to end (the FP64 oracle of the metric).
* ``np_*`` functions are independent vectorised numpy twins for tiny inputs,
  used to cross-check the C code.

Pinning (what checks each function; see tests/test_oracle_*.py):
* cubic_weights/sample1/sample2: closed forms (cubic reproduction, partition of
  unity, boundary rules of ffbp.py:27,35) and brute force against the twins.
* backproject_points/bp_polar/form_ffbp/evaluate_image: the reference's
  known-answer test fast_bp/validate.py (exact counts and shapes, FFBP-vs-direct
  errors in fast_bp/results/synthetic_validation.json) and the argmax pixels
  of the seeded operator validations (gpu_baseline.py:121-227,
  resident_experiment.py:146-183, resident_followup.py:48-80), plus the
  closed-form focused phase of a single analytic point target.
* regrid: identity invariant on identical grids plus the FFBP KAT above.
* target_contributions: identity sum_n w_n*phase_n*contrib == backproject_points.
* linear and kaiser interpolators (no reference counterpart): exact
  reproduction of linear / band-limited sequences and their definitions.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"

LAGRANGE4, LINEAR, KAISER8 = 0, 1, 2
INTERP = {"lagrange4": LAGRANGE4, "linear": LINEAR, "kaiser8": KAISER8}
C = 299792458.0

_lib = None


def build():
    """Compile liborc.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_l = ctypes.c_long
_d = ctypes.c_double
_i = ctypes.c_int


def _declare(L):
    L.orc_num_threads.restype = _i
    L.orc_cubic_weights.argtypes = [_d, _dp]
    L.orc_kaiser_sinc.argtypes = [_d]
    L.orc_kaiser_sinc.restype = _d
    L.orc_kaiser_weights.argtypes = [_d, _dp]
    L.orc_kaiser_table_value.argtypes = [_i]
    L.orc_kaiser_table_value.restype = _d
    L.orc_sample1.argtypes = [_vp, _l, _d, _i, _dp]
    L.orc_sample2.argtypes = [_vp, _i, _l, _l, _d, _d, _i, _dp]
    L.orc_backproject_points.argtypes = [_vp, _l, _l, _vp, _vp, _l, _d, _d, _d, _vp, _vp, _i, _vp]
    L.orc_bp_polar.argtypes = [_vp, _l, _l, _vp, _vp, _d, _d, _l, _d, _d, _l, _d, _d, _d,
                               _vp, _vp, _i, _i, _vp]
    L.orc_regrid.argtypes = [_vp, _i, _l, _l, _vp, _d, _d, _d, _d, _vp, _d, _d, _l, _d, _d, _l,
                             _d, _i, _i, _vp]
    L.orc_bp_polar_blocks.argtypes = [_vp, _l, _l, _vp, _vp, _d, _d, _l, _d, _d, _l, _d, _d, _d,
                                      _vp, _vp, _i, _i, _vp, _l, _i, _i, _vp]
    L.orc_regrid_blocks.argtypes = [_vp, _i, _l, _l, _vp, _d, _d, _d, _d, _vp, _d, _d, _l, _d, _d,
                                    _l, _d, _i, _i, _vp, _l, _i, _i, _i, _vp]
    L.orc_mark_support.argtypes = [_vp, _l, _l, _vp, _l, _l, _i, _vp, _l, _i, _i, _i, _i, _vp]
    L.orc_evaluate_image.argtypes = [_vp, _i, _l, _l, _vp, _d, _d, _d, _d, _vp, _l, _d, _i, _vp]
    L.orc_target_contributions.argtypes = [_vp, _l, _l, _vp, _vp, _l, _d, _d, _d, _i, _vp]
    L.orc_backproject_points_adjoint.argtypes = [_vp, _l, _l, _vp, _vp, _l, _d, _d, _d, _vp, _vp,
                                                 _i, _vp, _vp, _vp]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _interp(interp):
    return INTERP[interp] if isinstance(interp, str) else int(interp)


def _c64(data):
    data = np.ascontiguousarray(data)
    if data.dtype != np.complex64:
        data = data.astype(np.complex64)
    return data


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


def num_threads():
    return int(lib().orc_num_threads())


# ---------------------------------------------------------------------------
# Interpolation primitives
# ---------------------------------------------------------------------------

def cubic_weights(t):
    """fast_bp/ffbp.py:17-21."""
    w = np.zeros(4)
    lib().orc_cubic_weights(float(t), w.ctypes.data_as(_dp))
    return tuple(float(x) for x in w)


def kaiser_weights(t):
    w = np.zeros(8)
    lib().orc_kaiser_weights(float(t), w.ctypes.data_as(_dp))
    return w


def sample1(row, q, interp="lagrange4"):
    """fast_bp/ffbp.py:24-29 on a complex64 row."""
    row = _c64(row)
    out = np.zeros(2)
    lib().orc_sample1(_ptr(row), len(row), float(q), _interp(interp), out.ctypes.data_as(_dp))
    return complex(out[0], out[1])


def sample2(img, r, u, interp="lagrange4"):
    """fast_bp/ffbp.py:32-40 on a complex64 or complex128 [nr, nu] image."""
    img = np.ascontiguousarray(img)
    is128 = int(img.dtype == np.complex128)
    if not is128:
        img = _c64(img)
    out = np.zeros(2)
    lib().orc_sample2(_ptr(img), is128, img.shape[0], img.shape[1], float(r), float(u),
                      _interp(interp), out.ctypes.data_as(_dp))
    return complex(out[0], out[1])


# ---------------------------------------------------------------------------
# Kernels
# ---------------------------------------------------------------------------

def backproject_points(data, pos, xyz, r_start, dr, k, weight, phase, interp="lagrange4"):
    """fast_bp/ffbp.py:43-55; returns complex128 (round for the c64 policy)."""
    data = _c64(data)
    P, S = data.shape
    pos = _f64(pos, (P, 3))
    xyz = _f64(xyz)
    xyz = xyz.reshape(-1, 3)
    weight = _f64(weight, (P,))
    phase = _f64(phase, (P,))
    out = np.zeros(len(xyz), np.complex128)
    lib().orc_backproject_points(_ptr(data), P, S, _ptr(pos), _ptr(xyz), len(xyz),
                                 float(r_start), float(dr), float(k), _ptr(weight),
                                 _ptr(phase), _interp(interp), _ptr(out))
    return out


def bp_polar(data, pos, origin, r0, dr, nr, u0, du, nu, r_start, data_dr, k, weight, phase,
             side, interp="lagrange4"):
    """fast_bp/ffbp.py:58-76; returns complex128 [nr, nu]."""
    data = _c64(data)
    P, S = data.shape
    pos = _f64(pos, (P, 3))
    origin = _f64(origin, (3,))
    weight = _f64(weight, (P,))
    phase = _f64(phase, (P,))
    out = np.zeros((int(nr), int(nu)), np.complex128)
    lib().orc_bp_polar(_ptr(data), P, S, _ptr(pos), _ptr(origin), float(r0), float(dr), int(nr),
                       float(u0), float(du), int(nu), float(r_start), float(data_dr), float(k),
                       _ptr(weight), _ptr(phase), int(side), _interp(interp), _ptr(out))
    return out


def regrid(img, src_origin, sr0, sdr, su0, sdu, dst_origin, r0, dr, nr, u0, du, nu, k, side,
           interp="lagrange4"):
    """fast_bp/ffbp.py:79-95; img complex64 or complex128; returns complex128."""
    img = np.ascontiguousarray(img)
    is128 = int(img.dtype == np.complex128)
    if not is128:
        img = _c64(img)
    src_origin = _f64(src_origin, (3,))
    dst_origin = _f64(dst_origin, (3,))
    out = np.zeros((int(nr), int(nu)), np.complex128)
    lib().orc_regrid(_ptr(img), is128, img.shape[0], img.shape[1], _ptr(src_origin), float(sr0),
                     float(sdr), float(su0), float(sdu), _ptr(dst_origin), float(r0), float(dr),
                     int(nr), float(u0), float(du), int(nu), float(k), int(side),
                     _interp(interp), _ptr(out))
    return out


def evaluate_image(img, origin, r0, dr, u0, du, xyz, k, interp="lagrange4"):
    """fast_bp/ffbp.py:189-198; returns complex128 [M]."""
    img = np.ascontiguousarray(img)
    is128 = int(img.dtype == np.complex128)
    if not is128:
        img = _c64(img)
    origin = _f64(origin, (3,))
    xyz = _f64(xyz).reshape(-1, 3)
    out = np.zeros(len(xyz), np.complex128)
    lib().orc_evaluate_image(_ptr(img), is128, img.shape[0], img.shape[1], _ptr(origin),
                             float(r0), float(dr), float(u0), float(du), _ptr(xyz), len(xyz),
                             float(k), _interp(interp), _ptr(out))
    return out


def target_contributions(data, pos, xyz, r_start, dr, k, interp="lagrange4"):
    """fast_bp/ffbp.py:201-211; returns complex128 [M, P]."""
    data = _c64(data)
    P, S = data.shape
    pos = _f64(pos, (P, 3))
    xyz = _f64(xyz).reshape(-1, 3)
    out = np.zeros((len(xyz), P), np.complex128)
    lib().orc_target_contributions(_ptr(data), P, S, _ptr(pos), _ptr(xyz), len(xyz),
                                   float(r_start), float(dr), float(k), _interp(interp),
                                   _ptr(out))
    return out


def backproject_points_adjoint(data, pos, xyz, r_start, dr, k, weight, phase, grad,
                               interp="lagrange4"):
    """FP64 adjoint of backproject_points: (dL/dphase [P], dL/dpos [P, 3]) for
    dL = Re sum conj(grad) dI.  Pinned by central finite differences."""
    data = _c64(data)
    P, S = data.shape
    pos = _f64(pos, (P, 3))
    xyz = _f64(xyz).reshape(-1, 3)
    weight = _f64(weight, (P,))
    phase = _f64(phase, (P,))
    grad = np.ascontiguousarray(grad, dtype=np.complex128).reshape(len(xyz))
    gphi = np.zeros(P)
    gpos = np.zeros((P, 3))
    lib().orc_backproject_points_adjoint(_ptr(data), P, S, _ptr(pos), _ptr(xyz), len(xyz),
                                         float(r_start), float(dr), float(k), _ptr(weight),
                                         _ptr(phase), _interp(interp), _ptr(grad), _ptr(gphi),
                                         _ptr(gpos))
    return gphi, gpos


def entropy_and_grad(image):
    """Image entropy E = -sum p ln p, p = |I|^2 / sum |I|^2, and its gradient
    g = dE/dI in the convention dE = Re sum conj(g) dI, i.e.
    g = -2 (ln p + E) I / S (closed form)."""
    I = np.asarray(image, np.complex128)
    a2 = np.abs(I) ** 2
    S = a2.sum()
    p = a2 / S
    lp = np.log(np.where(p > 0, p, 1.0))
    E = float(-(p * lp).sum())
    g = -2.0 * (lp + E) * I / S
    return E, g


def polar_points(origin, r, u, side=1):
    """fast_bp/ffbp.py:183-186."""
    rr, uu = np.meshgrid(r, u, indexing="ij")
    gr = np.sqrt(rr ** 2 - origin[2] ** 2)
    return np.column_stack([(origin[0] + side * gr * np.sqrt(1 - uu ** 2)).ravel(),
                            (origin[1] + gr * uu).ravel(), np.zeros(rr.size)])


# ---------------------------------------------------------------------------
# Recursive FFBP (fast_bp/ffbp.py:114-180)
# ---------------------------------------------------------------------------

@dataclass
class PolarImage:
    """fast_bp/ffbp.py:98-111."""
    image: np.ndarray
    origin: np.ndarray
    r0: float
    dr: float
    u0: float
    du: float
    side: int

    @property
    def r(self):
        return self.r0 + self.dr * np.arange(self.image.shape[0])

    @property
    def u(self):
        return self.u0 + self.du * np.arange(self.image.shape[1])


def _node_grid_rule(pos, extent, side, wavelength, range_spacing, angular_oversample):
    """The node polar grid of fast_bp/ffbp.py:126-152 as a function of the
    pulse span [a, b): (origin, r0, nr, u0, du, nu).  The sector boundary is
    sampled at 129 points per edge, seen from the node's mean position."""
    root_origin = pos.mean(axis=0)
    rmin, rmax, umin, umax = extent
    us = np.linspace(umin, umax, 129)
    rs = np.linspace(rmin, rmax, 129)
    rb = np.r_[np.full_like(us, rmin), np.full_like(us, rmax), rs, rs]
    ub = np.r_[us, us, np.full_like(rs, umin), np.full_like(rs, umax)]
    g = np.sqrt(rb ** 2 - root_origin[2] ** 2)
    bx = root_origin[0] + side * g * np.sqrt(1 - ub ** 2)
    by = root_origin[1] + g * ub

    def grid_of(a, b):
        origin = pos[a:b].mean(axis=0)
        dx = bx - origin[0]
        dy = by - origin[1]
        gr = np.sqrt(dx * dx + dy * dy)
        rr = np.sqrt(gr * gr + origin[2] ** 2)
        uu = dy / gr
        aperture = np.linalg.norm(np.ptp(pos[a:b, :2], axis=0))
        du = min(.025, wavelength / (2 * max(aperture, .1) * angular_oversample))
        u0 = max(-.995, float(uu.min()) - 8 * du)
        u1 = min(.995, float(uu.max()) + 8 * du)
        r0 = float(rr.min()) - 8 * range_spacing
        nr = int(np.ceil((rr.max() - r0) / range_spacing)) + 9
        nu = int(np.ceil((u1 - u0) / du)) + 1
        return origin, r0, nr, u0, du, nu
    return grid_of


def form_ffbp(data, pos, r_start, data_dr, fc, extent, side=1, leaf_size=128, range_spacing=.25,
              angular_oversample=2.5, weight=None, phase=None, policy="parity",
              leaf_interp="lagrange4", merge_interp="lagrange4", collect_nodes=None):
    """Restatement of fast_bp/ffbp.py:114-180.

    ``policy="parity"`` rounds to complex64 where the reference does;
    ``policy="fp64"`` keeps complex128.  ``collect_nodes`` (a list) receives
    ``(a, b, origin, r0, nr, u0, du, nu)`` per node in DFS pre-order, i.e. the
    reference plan's node order (gpu_backend.py:154-191).
    """
    if policy not in ("parity", "fp64"):
        raise ValueError("policy must be parity or fp64")
    start = time.perf_counter()
    k = 4 * np.pi * fc / C
    wavelength = C / abs(fc)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    data = _c64(data)
    if weight is None:
        weight = np.ones(len(pos))
    if phase is None:
        phase = np.zeros(len(pos))
    weight = np.ascontiguousarray(weight, dtype=np.float64)
    phase = np.ascontiguousarray(phase, dtype=np.float64)
    store = np.complex64 if policy == "parity" else np.complex128
    rmin, rmax, umin, umax = extent
    grid_of = _node_grid_rule(pos, extent, side, wavelength, range_spacing, angular_oversample)
    counts = {"leaves": 0, "merges": 0, "leaf_pixel_pulse_products": 0, "merged_pixels": 0}

    def node(a, b):
        origin, r0, nr, u0, du, nu = grid_of(a, b)
        if collect_nodes is not None:
            collect_nodes.append((a, b, origin.copy(), r0, nr, u0, du, nu))
        if b - a <= leaf_size:
            im = bp_polar(data[a:b], pos[a:b], origin, r0, range_spacing, nr, u0, du, nu,
                          r_start, data_dr, k, weight[a:b], phase[a:b], side,
                          interp=leaf_interp).astype(store)
            counts["leaves"] += 1
            counts["leaf_pixel_pulse_products"] += nr * nu * (b - a)
        else:
            mid = (a + b) // 2
            left = node(a, mid)
            im = regrid(left.image, left.origin, left.r0, left.dr, left.u0, left.du, origin, r0,
                        range_spacing, nr, u0, du, nu, k, side, interp=merge_interp).astype(store)
            del left
            right = node(mid, b)
            im += regrid(right.image, right.origin, right.r0, right.dr, right.u0, right.du,
                         origin, r0, range_spacing, nr, u0, du, nu, k, side,
                         interp=merge_interp).astype(store)
            del right
            counts["merges"] += 1
            counts["merged_pixels"] += 2 * nr * nu
        return PolarImage(im, origin, r0, range_spacing, u0, du, side)

    result = node(0, len(pos))
    ri = np.flatnonzero((result.r >= rmin) & (result.r <= rmax))
    ui = np.flatnonzero((result.u >= umin) & (result.u <= umax))
    result.r0 += ri[0] * result.dr
    result.u0 += ui[0] * result.du
    result.image = result.image[ri[0]:ri[-1] + 1, ui[0]:ui[-1] + 1].copy()
    counts["elapsed_s"] = time.perf_counter() - start
    counts["shape"] = list(result.image.shape)
    return result, counts


# Stencil of each interpolator: first tap offset from floor(q), tap count
# (bp_oracle.c: sample1/sample2).
STENCIL = {"lagrange4": (-1, 4), "linear": (0, 2), "kaiser8": (-3, 8)}


def _mark_stencils(mask, nr, nu, qr, qu, interp, bs, margin=1):
    """Set the blocks of ``mask`` (bs x bs pixels per entry) that the stencils
    at fractional positions (qr, qu), widened by ``margin`` cells, touch."""
    first, taps = STENCIL[interp]
    ok = np.isfinite(qr) & np.isfinite(qu)
    fr, fu = np.floor(qr[ok]), np.floor(qu[ok])
    lo_r, hi_r = fr + first - margin, fr + first + taps - 1 + margin
    lo_u, hi_u = fu + first - margin, fu + first + taps - 1 + margin
    keep = (hi_r >= 0) & (lo_r <= nr - 1) & (hi_u >= 0) & (lo_u <= nu - 1)
    br0 = (np.clip(lo_r[keep], 0, nr - 1) // bs).astype(np.int64)
    br1 = (np.clip(hi_r[keep], 0, nr - 1) // bs).astype(np.int64)
    bu0 = (np.clip(lo_u[keep], 0, nu - 1) // bs).astype(np.int64)
    bu1 = (np.clip(hi_u[keep], 0, nu - 1) // bs).astype(np.int64)
    span = (taps + 2 * margin + bs - 2) // bs + 1   # blocks a box can cross per axis
    for a in range(span):
        for b in range(span):
            mask[np.minimum(br0 + a, br1), np.minimum(bu0 + b, bu1)] = True


def ffbp_at_points(data, pos, r_start, data_dr, fc, extent, xyz, side=1, leaf_size=128,
                   range_spacing=.25, angular_oversample=2.5, weight=None, phase=None,
                   policy="fp64", leaf_interp="lagrange4", merge_interp="lagrange4",
                   eval_interp=None, block=8):
    """``evaluate_image(form_ffbp(...), xyz)`` computing only what those
    values depend on.

    Same tree, node grids and per-pixel arithmetic as ``form_ffbp`` (the C
    block kernels share the pixel bodies of bp_polar/regrid), so every value
    equals the full computation's bit for bit; the saving is that a node
    pixel is computed only when it lies in a block of ``block`` x ``block``
    pixels that some stencil of a required parent pixel (or of an output
    point, at the root crop) touches, each stencil widened by one cell on
    every side.  Pixels outside those blocks stay 0 and are never read by a
    computed pixel.  Returns (values complex128 [M], info dict)."""
    if policy not in ("parity", "fp64"):
        raise ValueError("policy must be parity or fp64")
    eval_interp = merge_interp if eval_interp is None else eval_interp
    start = time.perf_counter()
    k = 4 * np.pi * fc / C
    wavelength = C / abs(fc)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    data = _c64(data)
    weight = np.ones(len(pos)) if weight is None else weight
    phase = np.zeros(len(pos)) if phase is None else phase
    weight = np.ascontiguousarray(weight, dtype=np.float64)
    phase = np.ascontiguousarray(phase, dtype=np.float64)
    store = np.complex64 if policy == "parity" else np.complex128
    rmin, rmax, umin, umax = extent
    xyz = _f64(xyz).reshape(-1, 3)
    grid_of = _node_grid_rule(pos, extent, side, wavelength, range_spacing, angular_oversample)

    # the tree in pre-order: (a, b, geometry, nr, nu, children)
    nodes = []

    def build(a, b):
        origin, r0, nr, u0, du, nu = grid_of(a, b)
        i = len(nodes)
        nodes.append({"a": a, "b": b, "geo": (origin, r0, range_spacing, u0, du), "nr": nr,
                      "nu": nu, "kids": ()})
        if b - a > leaf_size:
            mid = (a + b) // 2
            nodes[i]["kids"] = (build(a, mid), build(mid, b))
        return i
    build(0, len(pos))

    def blocks_of(n):
        return np.zeros((-(-n["nr"] // block), -(-n["nu"] // block)), bool)

    # root crop (ffbp.py:172-177) and the root blocks the evaluation reads
    root = nodes[0]
    origin, r0, dr, u0, du = root["geo"]
    rr_ax = r0 + dr * np.arange(root["nr"])
    uu_ax = u0 + du * np.arange(root["nu"])
    ri = np.flatnonzero((rr_ax >= rmin) & (rr_ax <= rmax))
    ui = np.flatnonzero((uu_ax >= umin) & (uu_ax <= umax))
    crop = (int(ri[0]), int(ri[-1]) + 1, int(ui[0]), int(ui[-1]) + 1)
    cr0, cu0 = r0 + ri[0] * dr, u0 + ui[0] * du
    dxp = xyz[:, 0] - origin[0]
    dyp = xyz[:, 1] - origin[1]
    gp = np.sqrt(dxp * dxp + dyp * dyp)
    with np.errstate(divide="ignore", invalid="ignore"):
        gp = np.where(gp == 0.0, np.nan, gp)
        qr = (np.sqrt(gp * gp + origin[2] ** 2) - cr0) / dr + crop[0]
        qu = (dyp / gp - cu0) / du + crop[2]
    root["mask"] = blocks_of(root)
    _mark_stencils(root["mask"], root["nr"], root["nu"], qr, qu, eval_interp, block)
    # parents before children: each child's blocks from its parent's
    first, taps = STENCIL[merge_interp]
    for n in nodes:
        blocks = np.ascontiguousarray(np.argwhere(n["mask"]), dtype=np.int32)
        n["blocks"] = blocks
        pgeo = np.r_[n["geo"][0], n["geo"][1:]].astype(np.float64)
        for c in n["kids"]:
            child = nodes[c]
            child["mask"] = blocks_of(child)
            cgeo = np.r_[child["geo"][0], child["geo"][1:]].astype(np.float64)
            lib().orc_mark_support(_ptr(pgeo), n["nr"], n["nu"], _ptr(cgeo), child["nr"],
                                   child["nu"], int(side), _ptr(blocks), len(blocks), block,
                                   first, taps, 1, _ptr(child["mask"]))
    info = {"node_pixels": 0, "computed_pixels": 0, "nodes": len(nodes)}
    out_c64 = int(store == np.complex64)

    def compute(i):
        n = nodes[i]
        origin, r0, dr, u0, du = n["geo"]
        blocks = n["blocks"]
        info["node_pixels"] += n["nr"] * n["nu"]
        info["computed_pixels"] += int(min(len(blocks) * block * block, n["nr"] * n["nu"]))
        out = np.zeros((n["nr"], n["nu"]), store)
        if not n["kids"]:
            a, b = n["a"], n["b"]
            d = np.ascontiguousarray(data[a:b])
            lib().orc_bp_polar_blocks(
                _ptr(d), b - a, d.shape[1], _ptr(np.ascontiguousarray(pos[a:b])),
                _ptr(np.ascontiguousarray(origin)), float(r0), float(dr), n["nr"], float(u0),
                float(du), n["nu"], float(r_start), float(data_dr), float(k),
                _ptr(np.ascontiguousarray(weight[a:b])), _ptr(np.ascontiguousarray(phase[a:b])),
                int(side), _interp(leaf_interp), _ptr(blocks), len(blocks), block, out_c64,
                _ptr(out))
            return out
        # ffbp.py:162-167: the first child's regrid stored, the second's added,
        # both in the storage type
        for j, c in enumerate(n["kids"]):
            child_img = compute(c)
            co, sr0, sdr, su0, sdu = nodes[c]["geo"]
            lib().orc_regrid_blocks(
                _ptr(child_img), int(child_img.dtype == np.complex128), child_img.shape[0],
                child_img.shape[1], _ptr(np.ascontiguousarray(co)), float(sr0), float(sdr),
                float(su0), float(sdu), _ptr(np.ascontiguousarray(origin)), float(r0), float(dr),
                n["nr"], float(u0), float(du), n["nu"], float(k), int(side),
                _interp(merge_interp), _ptr(blocks), len(blocks), block, out_c64, j, _ptr(out))
            del child_img
        return out

    img = compute(0)
    cropped = np.ascontiguousarray(img[crop[0]:crop[1], crop[2]:crop[3]])
    del img
    values = evaluate_image(cropped, origin, cr0, dr, cu0, du, xyz, k, interp=eval_interp)
    info["computed_fraction"] = info["computed_pixels"] / info["node_pixels"]
    info["crop_shape"] = [crop[1] - crop[0], crop[3] - crop[2]]
    info["elapsed_s"] = time.perf_counter() - start
    return values, info


# ---------------------------------------------------------------------------
# Independent numpy twins (tiny inputs only)
# ---------------------------------------------------------------------------

def np_cubic_weights(t):
    t = np.asarray(t, dtype=np.float64)
    return np.stack([-t * (t - 1) * (t - 2) / 6, (t + 1) * (t - 1) * (t - 2) / 2,
                     -(t + 1) * t * (t - 2) / 2, (t + 1) * t * (t - 1) / 6], axis=-1)


def np_sample1(row, q):
    """Vectorised sample1 over an array of coordinates q."""
    row = np.asarray(row).astype(np.complex128)
    q = np.asarray(q, dtype=np.float64)
    i = np.floor(q).astype(np.int64)
    ok = (i >= 1) & (i + 2 < len(row))
    ic = np.clip(i, 1, max(1, len(row) - 3))
    w = np_cubic_weights(q - i)
    taps = np.stack([row[np.clip(ic + a - 1, 0, len(row) - 1)] for a in range(4)], axis=-1)
    return np.where(ok, np.sum(w * taps, axis=-1), 0)


def np_bp_polar(data, pos, origin, r0, dr, nr, u0, du, nu, r_start, data_dr, k, weight, phase,
                side):
    r = r0 + dr * np.arange(nr)
    u = u0 + du * np.arange(nu)
    rr, uu = np.meshgrid(r, u, indexing="ij")
    ground = np.sqrt(np.maximum(0., rr * rr - origin[2] ** 2))
    x = origin[0] + side * ground * np.sqrt(np.maximum(0., 1 - uu * uu))
    y = origin[1] + ground * uu
    total = np.zeros(rr.shape, np.complex128)
    for n in range(len(pos)):
        d = np.sqrt((x - pos[n, 0]) ** 2 + (y - pos[n, 1]) ** 2 + pos[n, 2] ** 2)
        v = np_sample1(data[n], (d - r_start) / data_dr)
        total += weight[n] * v * np.exp(1j * (k * (d - rr) + phase[n]))
    return total


def np_backproject_points(data, pos, xyz, r_start, dr, k, weight, phase):
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    total = np.zeros(len(xyz), np.complex128)
    for n in range(len(pos)):
        d = np.linalg.norm(xyz - pos[n], axis=1)
        v = np_sample1(data[n], (d - r_start) / dr)
        total += weight[n] * v * np.exp(1j * (k * d + phase[n]))
    return total


def np_sample2(img, qr, qu):
    img = np.asarray(img).astype(np.complex128)
    qr = np.asarray(qr, dtype=np.float64)
    qu = np.asarray(qu, dtype=np.float64)
    ir = np.floor(qr).astype(np.int64)
    iu = np.floor(qu).astype(np.int64)
    nr, nu = img.shape
    ok = (ir >= 1) & (iu >= 1) & (ir + 2 < nr) & (iu + 2 < nu)
    wr = np_cubic_weights(qr - ir)
    wu = np_cubic_weights(qu - iu)
    z = np.zeros(qr.shape, np.complex128)
    for a in range(4):
        for b in range(4):
            z += img[np.clip(ir + a - 1, 0, nr - 1), np.clip(iu + b - 1, 0, nu - 1)] * wr[..., a] * wu[..., b]
    return np.where(ok, z, 0)


def np_regrid(img, src_origin, sr0, sdr, su0, sdu, dst_origin, r0, dr, nr, u0, du, nu, k, side):
    r = r0 + dr * np.arange(nr)
    u = u0 + du * np.arange(nu)
    rr_, uu_ = np.meshgrid(r, u, indexing="ij")
    gr = np.sqrt(np.maximum(0., rr_ * rr_ - dst_origin[2] ** 2))
    x = dst_origin[0] + side * gr * np.sqrt(np.maximum(0., 1 - uu_ * uu_))
    y = dst_origin[1] + gr * uu_
    dx = x - src_origin[0]
    dy = y - src_origin[1]
    g = np.sqrt(dx * dx + dy * dy)
    rr = np.sqrt(g * g + src_origin[2] ** 2)
    uu = dy / g
    v = np_sample2(img, (rr - sr0) / sdr, (uu - su0) / sdu)
    return v * np.exp(1j * k * (rr - rr_))


# ---------------------------------------------------------------------------
# Range compression (the step before BP), FP64 restatement of
# fast_bp/sar_io.py:116-139.  Pinned by the reference's fractional-delay
# known answer (fast_bp/validate.py:55-59: a 1 us, 200 MHz chirp centred at
# sample 1512.4 compresses, at 4x Fourier oversampling, to a peak at 1512.5 --
# results/synthetic_validation.json ``lfm_peak_sample``) and by closed forms
# (a chirp centred on an integer sample peaks there at the compressed energy).
# ---------------------------------------------------------------------------

FS_REF = 400e6


def lfm_reference(tp, bandwidth, chirp_sign, fs=FS_REF):
    """Centred LFM reference with a periodic Hamming window (sar_io.py:124-128),
    complex128: samples at (n - floor(L/2)) / fs, L = round(tp fs)."""
    from scipy.signal.windows import hamming
    length = int(round(tp * fs))
    tr = (np.arange(length) - length // 2) / fs
    ref = np.exp(1j * np.pi * chirp_sign * bandwidth / tp * tr ** 2)
    return ref * hamming(length, sym=False)


def range_compress(raw, tp=1e-6, bandwidth=200e6, chirp_sign=-1, oversample=1, fs=FS_REF,
                   nfft=None):
    """Linear correlation with the centred reference via zero-padded FFTs
    (sar_io.py:129-139), complex128.  Sample k of the output is the received
    pulse centred at raw sample k; ``oversample`` > 1 Fourier-interpolates the
    length-nfft correlation (scipy.signal.resample semantics, the +N/2 bin
    split between +N/2 and -N/2) before the same crop."""
    from scipy import fft, signal
    raw = np.asarray(raw, np.complex128)
    ref = lfm_reference(tp, bandwidth, chirp_sign, fs)
    length = len(ref)
    if nfft is None:
        nfft = fft.next_fast_len(raw.shape[1] + length - 1)
    h = np.conj(ref[::-1])
    spectrum = fft.fft(raw - raw.mean(axis=1, keepdims=True), nfft, axis=1)
    spectrum *= fft.fft(h, nfft)[None, :]
    offset = length - 1 - length // 2
    corr = fft.ifft(spectrum, axis=1)
    if oversample == 1:
        return corr[:, offset:offset + raw.shape[1]].copy()
    corr = signal.resample(corr, nfft * oversample, axis=1)
    return corr[:, offset * oversample:(offset + raw.shape[1]) * oversample].copy()


def lfm_echo(samples, centre, tp=1e-6, bandwidth=200e6, fs=FS_REF):
    """Baseband LFM echo of the validate.py KAT (validate.py:55-56): the
    down-chirp exp(-j pi B/tp t^2), |t| < tp/2, centred at fractional sample
    ``centre``; complex64 [1, samples]."""
    tr = (np.arange(samples) - centre) / fs
    x = np.exp(-1j * np.pi * bandwidth / tp * tr ** 2) * (np.abs(tr) < tp / 2)
    return x.astype(np.complex64)[None, :]


def decode_frames(path, offset, frame_size, nr, frame_index):
    """I/Q of capture frames (fast_bp/sar_io.py:79-81 Capture.read) read with
    plain byte arithmetic: frame f starts at offset + f * frame_size, its
    int16 I/Q pairs at +160; complex64 [len(frame_index), nr]."""
    data = np.fromfile(path, dtype=np.uint8)
    out = np.zeros((len(frame_index), nr), np.complex64)
    for row, f in enumerate(frame_index):
        start = offset + int(f) * frame_size + 160
        iq = data[start:start + 4 * nr].view("<i2").reshape(nr, 2).astype(np.float32)
        out[row] = iq[:, 0] + 1j * iq[:, 1]
    return out
