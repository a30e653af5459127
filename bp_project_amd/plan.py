"""FP64 host planning of the FFBP tree (no device code, no oracle).

``build_ffbp_plan`` reproduces ``fast_bp/gpu_backend.py:121-224``: the same
node tree as the recursive driver ``fast_bp/ffbp.py:114-180`` (origin = mean
antenna position, du rule, 8-cell guards, ``nr = ceil(...) + 9``, u clamped to
+-0.995, leaf when ``b - a <= leaf_size``, split at ``(a + b) // 2``), node
indices in depth-first pre-order and the root crop to the requested extent.
Every precision policy shares this plan, so FP32 never decides a ceil, a
guard or a crop.  The numpy expressions deliberately follow the reference's
so the plan is bit-identical to it.

``FlatPlan`` is the MI355X-side view of a plan: node geometry packed into
arrays (one row per node), nodes grouped by subtree height for level-batched
launches, and a single image arena with per-node offsets.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

C = 299792458.0


@dataclass
class NodePlan:
    """One FFBP tree node (fast_bp/gpu_backend.py:60-78)."""
    index: int
    a: int
    b: int
    origin: np.ndarray
    r0: float
    dr: float
    nr: int
    u0: float
    du: float
    nu: int
    side: int
    left: Optional["NodePlan"] = None
    right: Optional["NodePlan"] = None

    @property
    def is_leaf(self):
        return self.left is None


@dataclass
class FFBPPlan:
    """Frozen FFBP plan (fast_bp/gpu_backend.py:81-102)."""
    root: NodePlan
    origins: np.ndarray
    crop_r: tuple
    crop_u: tuple
    origin: np.ndarray
    r0: float
    dr: float
    nr: int
    u0: float
    du: float
    nu: int
    side: int
    counts: dict
    extent: tuple
    leaf_size: int
    angular_oversample: float

    @property
    def shape(self):
        return self.nr, self.nu


def _sector_boundary(root_origin, extent, side):
    rmin, rmax, umin, umax = (float(x) for x in extent)
    us = np.linspace(umin, umax, 129)
    rs = np.linspace(rmin, rmax, 129)
    rb = np.r_[np.full_like(us, rmin), np.full_like(us, rmax), rs, rs]
    ub = np.r_[us, us, np.full_like(rs, umin), np.full_like(rs, umax)]
    ground = np.sqrt(rb ** 2 - root_origin[2] ** 2)
    return root_origin[0] + side * ground * np.sqrt(1.0 - ub ** 2), root_origin[1] + ground * ub


def build_ffbp_plan(position, fc, extent, side=1, leaf_size=128, range_spacing=0.25,
                    angular_oversample=2.5):
    """Exact FP64 node tree and crop used by the recursive FFBP.

    Signature and results follow fast_bp/gpu_backend.py:121-224.
    """
    position = np.ascontiguousarray(position, dtype=np.float64)
    if position.ndim != 2 or position.shape[1] != 3 or len(position) == 0:
        raise ValueError("position must be a non-empty [pulse, 3] array")
    if int(leaf_size) <= 0:
        raise ValueError("leaf_size must be positive")
    wavelength = C / abs(float(fc))
    root_origin = position.mean(axis=0)
    rmin, rmax, umin, umax = (float(x) for x in extent)
    bx, by = _sector_boundary(root_origin, extent, side)
    counts = {"leaves": 0, "merges": 0, "leaf_pixel_pulse_products": 0, "merged_pixels": 0}
    nodes = []

    def make(a, b):
        origin = position[a:b].mean(axis=0)
        dx = bx - origin[0]
        dy = by - origin[1]
        gr = np.sqrt(dx * dx + dy * dy)
        rr = np.sqrt(gr * gr + origin[2] ** 2)
        uu = dy / gr
        aperture = np.linalg.norm(np.ptp(position[a:b, :2], axis=0))
        du = min(0.025, wavelength / (2.0 * max(aperture, 0.1) * angular_oversample))
        u0 = max(-0.995, float(uu.min()) - 8.0 * du)
        u1 = min(0.995, float(uu.max()) + 8.0 * du)
        r0 = float(rr.min()) - 8.0 * range_spacing
        nr = int(np.ceil((rr.max() - r0) / range_spacing)) + 9
        nu = int(np.ceil((u1 - u0) / du)) + 1
        node = NodePlan(index=len(nodes), a=a, b=b,
                        origin=np.ascontiguousarray(origin, dtype=np.float64),
                        r0=r0, dr=float(range_spacing), nr=nr, u0=u0, du=float(du), nu=nu,
                        side=int(side))
        nodes.append(node)
        if b - a <= leaf_size:
            counts["leaves"] += 1
            counts["leaf_pixel_pulse_products"] += nr * nu * (b - a)
        else:
            mid = (a + b) // 2
            node.left = make(a, mid)
            node.right = make(mid, b)
            counts["merges"] += 1
            counts["merged_pixels"] += 2 * nr * nu
        return node

    root = make(0, len(position))
    root_r = root.r0 + root.dr * np.arange(root.nr)
    root_u = root.u0 + root.du * np.arange(root.nu)
    ri = np.flatnonzero((root_r >= rmin) & (root_r <= rmax))
    ui = np.flatnonzero((root_u >= umin) & (root_u <= umax))
    if len(ri) == 0 or len(ui) == 0:
        raise ValueError("Requested extent does not intersect the planned root grid")
    origins = np.ascontiguousarray(np.stack([n.origin for n in nodes]), dtype=np.float64)
    counts = dict(counts)
    counts.update(nodes=len(nodes), shape=[int(len(ri)), int(len(ui))],
                  direct_pixel_pulse_products=int(len(ri) * len(ui) * len(position)))
    return FFBPPlan(root=root, origins=origins, crop_r=(int(ri[0]), int(ri[-1]) + 1),
                    crop_u=(int(ui[0]), int(ui[-1]) + 1), origin=root.origin.copy(),
                    r0=float(root.r0 + ri[0] * root.dr), dr=float(root.dr), nr=int(len(ri)),
                    u0=float(root.u0 + ui[0] * root.du), du=float(root.du), nu=int(len(ui)),
                    side=int(side), counts=counts, extent=tuple(float(x) for x in extent),
                    leaf_size=int(leaf_size), angular_oversample=float(angular_oversample))


def iter_nodes(root):
    """Depth-first pre-order, i.e. increasing ``index``."""
    stack = [root]
    while stack:
        node = stack.pop()
        yield node
        if not node.is_leaf:
            stack.append(node.right)
            stack.append(node.left)


# Column layout of FlatPlan.node_int (int32) and node_geo (float64).
NI_A, NI_B, NI_NR, NI_NU, NI_SIDE, NI_LEFT, NI_RIGHT, NI_HEIGHT = range(8)
NG_OX, NG_OY, NG_OZ, NG_R0, NG_DR, NG_U0, NG_DU = range(7)


@dataclass
class FlatPlan:
    """Array form of an FFBP plan for level-batched device execution.

    ``node_int[i]``  = (a, b, nr, nu, side, left, right, height), -1 for no child.
    ``node_geo[i]``  = (ox, oy, oz, r0, dr, u0, du) in FP64.
    ``offsets[i]``   = element offset of node i's complex64 image in the arena.
    ``levels[h]``    = node indices of subtree height h (0 = leaves), ascending.
    """
    node_int: np.ndarray
    node_geo: np.ndarray
    offsets: np.ndarray
    arena_elems: int
    levels: list = field(default_factory=list)

    @property
    def node_count(self):
        return len(self.node_int)


def flatten_plan(plan):
    """Pack ``plan`` for device execution.

    Every node gets its own slot of one arena (the reference's resident
    workspace policy: all schedule buffers preallocated, no liveness reuse,
    fast_bp/resident_backend.py:100).  Slots are 512-byte aligned.
    """
    nodes = sorted(iter_nodes(plan.root), key=lambda n: n.index)
    if [n.index for n in nodes] != list(range(len(nodes))):
        raise ValueError("plan node indices must be dense and start at zero")
    heights = {}

    def height(n):
        if n.index not in heights:
            heights[n.index] = 0 if n.is_leaf else 1 + max(height(n.left), height(n.right))
        return heights[n.index]

    for n in nodes:
        height(n)
    node_int = np.zeros((len(nodes), 8), np.int32)
    node_geo = np.zeros((len(nodes), 7), np.float64)
    for n in nodes:
        node_int[n.index] = (n.a, n.b, n.nr, n.nu, n.side,
                             -1 if n.is_leaf else n.left.index,
                             -1 if n.is_leaf else n.right.index, heights[n.index])
        node_geo[n.index] = (n.origin[0], n.origin[1], n.origin[2], n.r0, n.dr, n.u0, n.du)
    max_h = max(heights.values())
    levels = [[n.index for n in nodes if heights[n.index] == h] for h in range(max_h + 1)]
    sizes = node_int[:, NI_NR].astype(np.int64) * node_int[:, NI_NU].astype(np.int64)
    if np.any(sizes <= 0):
        raise ValueError("empty image grids are unsupported")
    offsets = np.zeros(len(nodes), np.int64)
    align = 64  # elements: 512-byte aligned node images
    cursor = 0
    for n in nodes:
        offsets[n.index] = cursor
        cursor += -(-int(sizes[n.index]) // align) * align
    return FlatPlan(node_int=node_int, node_geo=node_geo, offsets=offsets,
                    arena_elems=int(cursor), levels=levels)
