"""Gradient-driven autofocus: BP as a ``torch.autograd.Function`` with a
hand-written HIP adjoint.

The reference refines the trajectory by a phase-gradient estimate over
selected bright targets (fast_bp/autofocus.py:37-79, built on
``target_contributions``).  Here the image entropy
E = -sum p ln p, p = |I|^2 / sum |I|^2, is minimised directly over per-pulse
phase corrections and antenna positions with Adam; the gradients come from
``bpmi_bp_grid_adjoint`` (range and phase recomputed in the adjoint, nothing
stored per pixel and pulse).  PyTorch provides only the autograd plumbing,
the entropy expression and the optimiser; the image and its adjoint run in
libbpmi on the current HIP stream.

``pick_targets`` / ``estimate_update`` keep the reference's non-gradient
phase-gradient caller (fast_bp/autofocus.py:20-79) for comparison, with the
per-(target, pulse) contributions computed on the GPU.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib

_vp = ctypes.c_void_p


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the autograd BP needs a HIP device (torch.cuda)")
    return torch


def _stream():
    torch = _torch()
    return _vp(torch.cuda.current_stream().cuda_stream or None)


@dataclass
class GridOperator:
    """Fixed inputs of a Cartesian-grid BP: device data and weights, grid."""
    data: object            # torch complex64 [P, S] on cuda
    weight: object          # torch float64 [P] on cuda
    x0: float
    dx: float
    nx: int
    y0: float
    dy: float
    ny: int
    z0: float
    range_start: float
    data_dr: float
    k: float
    mode: str = "split"
    interp: str = "lagrange4"

    @property
    def pulses(self):
        return int(self.data.shape[0])

    @classmethod
    def from_scene(cls, scene, mode="split", interp="lagrange4"):
        torch = _torch()
        g = scene.grid
        return cls(data=torch.from_numpy(np.ascontiguousarray(scene.data)).cuda(),
                   weight=torch.from_numpy(np.ascontiguousarray(scene.weight, np.float64)).cuda(),
                   x0=g.x0, dx=g.dx, nx=g.nx, y0=g.y0, dy=g.dy, ny=g.ny, z0=g.z0,
                   range_start=scene.r_start, data_dr=scene.data_dr, k=scene.k, mode=mode,
                   interp=interp)

    def _forward_workspace(self):
        """The pulse-split partials of bpmi_bp_grid_ws for this grid (kept)."""
        ws = self.__dict__.get("_ws")
        if ws is None:
            torch = _torch()
            nbytes = ctypes.c_size_t()
            _lib.call("bpmi_bp_grid_workspace_bytes", self.nx, self.ny, self.pulses,
                      ctypes.byref(nbytes))
            ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=self.data.device)
            self.__dict__["_ws"] = ws
        return ws

    def image(self, phase, pos):
        """Forward BP (no autograd) into a new complex64 [ny, nx] tensor
        (pulse-split into chunks where the grid has few tiles)."""
        torch = _torch()
        out = torch.empty((self.ny, self.nx), dtype=torch.complex64, device=self.data.device)
        phase = phase.detach().to(torch.float64).contiguous()
        pos = pos.detach().to(torch.float64).contiguous()
        ws = self._forward_workspace()
        _lib.call("bpmi_bp_grid_ws", _vp(self.data.data_ptr()), int(self.data.shape[1]),
                  _vp(pos.data_ptr()), _vp(self.weight.data_ptr()), _vp(phase.data_ptr()), 0,
                  self.pulses, self.x0, self.dx, self.nx, self.y0, self.dy, self.ny, self.z0,
                  self.range_start, self.data_dr, self.k, _lib.mode_code(self.mode),
                  _lib.interp_code(self.interp), 0, _vp(out.data_ptr()), _vp(ws.data_ptr()),
                  int(ws.numel()), _stream())
        return out

    def image64(self, phase, pos):
        """FP64-mode forward BP with a complex128 output (finite differences)."""
        torch = _torch()
        out = torch.empty((self.ny, self.nx), dtype=torch.complex128, device=self.data.device)
        phase = phase.detach().to(torch.float64).contiguous()
        pos = pos.detach().to(torch.float64).contiguous()
        _lib.call("bpmi_bp_grid", _vp(self.data.data_ptr()), int(self.data.shape[1]),
                  _vp(pos.data_ptr()), _vp(self.weight.data_ptr()), _vp(phase.data_ptr()), 0,
                  self.pulses, self.x0, self.dx, self.nx, self.y0, self.dy, self.ny, self.z0,
                  self.range_start, self.data_dr, self.k, _lib.MODE_FP64,
                  _lib.interp_code(self.interp), _lib.OUT_C128, _vp(out.data_ptr()), _stream())
        return out

    def adjoint(self, phase, pos, grad_image):
        """(dL/dphase [P], dL/dpos [P, 3]) for dL = Re sum conj(grad) dI."""
        torch = _torch()
        P = self.pulses
        phase = phase.detach().to(torch.float64).contiguous()
        pos = pos.detach().to(torch.float64).contiguous()
        grad = grad_image.detach().to(torch.complex64).contiguous()
        nbytes = ctypes.c_size_t()
        _lib.call("bpmi_grid_adjoint_workspace_bytes", self.nx, self.ny, P, ctypes.byref(nbytes))
        work = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=self.data.device)
        gphi = torch.empty(P, dtype=torch.float64, device=self.data.device)
        gpos = torch.empty((P, 3), dtype=torch.float64, device=self.data.device)
        _lib.call("bpmi_bp_grid_adjoint", _vp(self.data.data_ptr()), int(self.data.shape[1]),
                  _vp(pos.data_ptr()), _vp(self.weight.data_ptr()), _vp(phase.data_ptr()), 0, P,
                  self.x0, self.dx, self.nx, self.y0, self.dy, self.ny, self.z0,
                  self.range_start, self.data_dr, self.k, _lib.interp_code(self.interp),
                  _vp(grad.data_ptr()), _vp(work.data_ptr()), nbytes.value,
                  _vp(gphi.data_ptr()), _vp(gpos.data_ptr()), _stream())
        return gphi, gpos


def _function():
    torch = _torch()

    class Backprojection(torch.autograd.Function):
        """I = BP(phase, pos); backward runs the HIP adjoint."""

        @staticmethod
        def forward(ctx, phase, pos, op):
            ctx.op = op
            ctx.save_for_backward(phase, pos)
            return op.image(phase, pos)

        @staticmethod
        def backward(ctx, grad_image):
            phase, pos = ctx.saved_tensors
            gphi, gpos = ctx.op.adjoint(phase, pos, grad_image)
            return gphi.to(phase.dtype), gpos.to(pos.dtype), None

    return Backprojection


_BP = None


def backproject(phase, pos, op):
    """Differentiable BP image (complex64 [ny, nx]) of per-pulse phase and positions."""
    global _BP
    if _BP is None:
        _BP = _function()
    return _BP.apply(phase, pos, op)


def entropy(image):
    """Image entropy in nats (differentiable)."""
    torch = _torch()
    a2 = image.real.double() ** 2 + image.imag.double() ** 2
    p = a2 / a2.sum()
    return -torch.special.xlogy(p, p).sum()


def gradient_autofocus(op, phase0, pos0, iterations=30, lr_phase=0.05, lr_pos=0.002,
                       optimise_pos=True, callback=None):
    """Adam on (phase correction, position correction) minimising entropy.

    Returns a dict with the corrected phase/pos (numpy), the entropy history
    and the per-iteration wall times (ms, synchronised).
    """
    torch = _torch()
    dev = op.data.device
    phase0 = torch.as_tensor(np.asarray(phase0, np.float64), device=dev)
    pos0 = torch.as_tensor(np.asarray(pos0, np.float64), device=dev)
    dphi = torch.zeros_like(phase0, requires_grad=True)
    dpos = torch.zeros_like(pos0, requires_grad=True)
    params = [{"params": [dphi], "lr": lr_phase}]
    if optimise_pos:
        params.append({"params": [dpos], "lr": lr_pos})
    opt = torch.optim.Adam(params)
    history, times = [], []
    best = (float("inf"), None, None)
    for it in range(iterations):
        torch.cuda.synchronize()
        t = time.perf_counter()
        opt.zero_grad()
        img = backproject(phase0 + dphi, pos0 + dpos, op)
        E = entropy(img)
        E.backward()
        opt.step()
        torch.cuda.synchronize()
        times.append(1000 * (time.perf_counter() - t))
        e = float(E.detach())
        history.append(e)
        if e < best[0]:
            best = (e, (phase0 + dphi).detach().cpu().numpy(), (pos0 + dpos).detach().cpu().numpy())
        if callback:
            callback(it, e)
    with torch.no_grad():
        final = float(entropy(op.image(phase0 + dphi, pos0 + dpos)))
    history.append(final)
    if final < best[0]:
        best = (final, (phase0 + dphi).detach().cpu().numpy(), (pos0 + dpos).detach().cpu().numpy())
    return {"phase": best[1], "pos": best[2], "entropy": history, "best_entropy": best[0],
            "ms_per_iteration": times}


def finite_difference_gradient(op, phase, pos, pulses, h=1e-4, mode="fp64"):
    """Central differences of the entropy w.r.t. phase and position of the
    given pulses (2 BP images per parameter).  ``mode="fp64"`` uses FP64 images
    with complex128 outputs (a 1e-4 rad change of one pulse moves a complex64
    pixel by only a few ulps); other modes use that mode's complex64 image."""
    torch = _torch()
    mode_saved = op.mode
    if mode and mode != "fp64":
        op.mode = mode
    try:
        phase = torch.as_tensor(np.asarray(phase, np.float64), device=op.data.device)
        pos = torch.as_tensor(np.asarray(pos, np.float64), device=op.data.device)

        def E(ph, ps):
            img = op.image64(ph, ps) if mode == "fp64" else op.image(ph, ps)
            return float(entropy(img))
        gphi = np.zeros(len(pulses))
        gpos = np.zeros((len(pulses), 3))
        for j, n in enumerate(pulses):
            a, b = phase.clone(), phase.clone()
            a[n] += h
            b[n] -= h
            gphi[j] = (E(a, pos) - E(b, pos)) / (2 * h)
            for axis in range(3):
                hp = h * 1e-2
                a, b = pos.clone(), pos.clone()
                a[n, axis] += hp
                b[n, axis] -= hp
                gpos[j, axis] = (E(phase, a) - E(phase, b)) / (2 * hp)
        return gphi, gpos
    finally:
        op.mode = mode_saved


def fd_autofocus(op, phase0, pos0, iterations=30, lr_phase=0.05, lr_pos=0.002, h=1e-4,
                 target=None):
    """The baseline that the adjoint replaces: the same Adam iteration as
    ``gradient_autofocus``, with the gradient taken by central differences of
    the entropy of FP64-mode BP images (complex128 outputs): two images per
    parameter, phase and x, y, z of every pulse, i.e. 8 P images per step
    (position steps h / 100, as ``finite_difference_gradient``).  The
    entropies stay on the device until the step's gradient is assembled.
    Stops after ``iterations`` steps or once the entropy of the current
    parameters is at most ``target``.  Returns the entropy history (before
    each step, then the final one), per-step wall times in ms (synchronised)
    and the number of images formed."""
    torch = _torch()
    dev = op.data.device
    phase0 = torch.as_tensor(np.asarray(phase0, np.float64), device=dev)
    pos0 = torch.as_tensor(np.asarray(pos0, np.float64), device=dev)
    P = op.pulses
    dphi = torch.zeros_like(phase0, requires_grad=True)
    dpos = torch.zeros_like(pos0, requires_grad=True)
    opt = torch.optim.Adam([{"params": [dphi], "lr": lr_phase},
                            {"params": [dpos], "lr": lr_pos}])
    hp = h * 1e-2
    history, times, images = [], [], 0

    def E(ph, ps):
        return entropy(op.image64(ph, ps))
    for _ in range(iterations):
        torch.cuda.synchronize()
        t = time.perf_counter()
        with torch.no_grad():
            phase = phase0 + dphi
            pos = pos0 + dpos
            e0 = float(E(phase, pos))
            images += 1
            if target is not None and e0 <= target:
                history.append(e0)
                break
            plus, minus = [], []
            for n in range(P):
                a, b = phase.clone(), phase.clone()
                a[n] += h
                b[n] -= h
                plus.append(E(a, pos))
                minus.append(E(b, pos))
                for axis in range(3):
                    a, b = pos.clone(), pos.clone()
                    a[n, axis] += hp
                    b[n, axis] -= hp
                    plus.append(E(phase, a))
                    minus.append(E(phase, b))
            images += 8 * P
            g = ((torch.stack(plus) - torch.stack(minus)).view(P, 4)
                 / torch.tensor([2 * h, 2 * hp, 2 * hp, 2 * hp], dtype=torch.float64,
                                device=dev))
        opt.zero_grad()
        dphi.grad = g[:, 0].contiguous()
        dpos.grad = g[:, 1:].contiguous()
        opt.step()
        torch.cuda.synchronize()
        times.append(1000 * (time.perf_counter() - t))
        history.append(e0)
    else:
        with torch.no_grad():
            history.append(float(E(phase0 + dphi, pos0 + dpos)))
            images += 1
    return {"entropy": history, "ms_per_iteration": times, "images": images,
            "phase": (phase0 + dphi).detach().cpu().numpy(),
            "pos": (pos0 + dpos).detach().cpu().numpy()}


def pick_targets(result, nr_tiles=12, nu_tiles=12):
    """Brightest polar pixel of each of nr_tiles x nu_tiles blocks of an FFBP
    root image, as ground points on z = 0, keeping those within 35 dB of the
    brightest (fast_bp/autofocus.py:20-34).  ``result`` is a ``PolarImage``.
    Returns (points [M, 3], powers [M])."""
    img = np.asarray(result.image)
    r_axis = np.asarray(result.r)
    u_axis = np.asarray(result.u)
    o = np.asarray(result.origin, np.float64)
    points, powers = [], []
    for rows in np.array_split(np.arange(len(r_axis)), nr_tiles):
        for cols in np.array_split(np.arange(len(u_axis)), nu_tiles):
            block = np.abs(img[rows[0]:rows[-1] + 1, cols[0]:cols[-1] + 1])
            i, j = np.unravel_index(block.argmax(), block.shape)
            r = r_axis[rows[i]]
            u = u_axis[cols[j]]
            ground = np.sqrt(r * r - o[2] ** 2)
            points.append([o[0] + result.side * ground * np.sqrt(1 - u * u), o[1] + ground * u, 0.])
            powers.append(float(block.max()) ** 2)
    powers = np.array(powers)
    points = np.array(points)
    keep = powers > powers.max() * 10 ** (-35 / 10)
    return points[keep], powers[keep]


def phase_gradient_update(contributions, pos, points, k, prf, cutoff):
    """Trajectory update from per-(target, pulse) contributions: the host
    half of fast_bp/autofocus.py:37-79.

    contributions: [M, P] complex (``target_contributions``); pos [P, 3];
    points [M, 3].  Steps, in the reference's order:

    1. zero-phase 4th-order Butterworth low-pass (cutoff Hz at the PRF) of
       each target's slow-time history;
    2. adjacent-pulse cross products -> phase gradient per target and pulse,
       with amplitude, per-target strength (capped at 3) and coherence
       weights, capped at 4;
    3. the phase gradient over k is the change of range: three iteratively
       reweighted least-squares solves of the 2 x 2 per-pulse system
       u_los . (dx, dy) = gradient / k (ridge 0.002 trace, robust scale
       1.4826 MAD, Huber-like reweighting at 2.5 scale);
    4. integrate, remove the linear trend, Savitzky-Golay smooth (window
       min(501, odd <= P/2*2-1), order 3) and damp to at most 1 m.

    Returns (delta [P, 3], stats dict) like the reference.
    """
    from scipy import signal
    xi = np.asarray(contributions)
    pos = np.asarray(pos, np.float64)
    points = np.asarray(points, np.float64).reshape(-1, 3)
    P = len(pos)
    sos = signal.butter(4, cutoff, fs=prf, output="sos")
    smooth = signal.sosfiltfilt(sos, xi, axis=1)
    cross = smooth[:, 1:] * np.conj(smooth[:, :-1])
    gradient = np.angle(cross)
    amp = np.abs(cross)
    target_power = np.mean(amp, axis=1)
    strength = np.minimum(np.sqrt(target_power / np.median(target_power)), 3.)
    coherence = np.abs(np.sum(cross, axis=1)) / np.sum(amp, axis=1)
    w = amp / (target_power[:, None] + 1e-12) * strength[:, None]
    w *= np.clip(coherence[:, None], 0, 1) ** 2
    w = np.minimum(w, 4.)
    los = points[:, None, :] - pos[None, 1:, :]
    dist = np.linalg.norm(los, axis=2)
    ux = los[:, :, 0] / dist
    uy = los[:, :, 1] / dist
    rhs = gradient / k
    dx = np.zeros(P - 1)
    dy = dx.copy()
    for _ in range(3):
        a = np.sum(w * ux * ux, axis=0)
        b = np.sum(w * ux * uy, axis=0)
        d = np.sum(w * uy * uy, axis=0)
        ex = np.sum(w * ux * rhs, axis=0)
        ey = np.sum(w * uy * rhs, axis=0)
        ridge = .002 * (a + d)
        det = (a + ridge) * (d + ridge) - b * b
        dx = (ex * (d + ridge) - ey * b) / np.maximum(det, 1e-12)
        dy = (ey * (a + ridge) - ex * b) / np.maximum(det, 1e-12)
        residual = rhs - ux * dx - uy * dy
        scale = 1.4826 * np.median(np.abs(residual), axis=0) + 1e-6
        w *= np.minimum(1, 2.5 * scale[None, :] / (np.abs(residual) + 1e-12))
    delta = np.column_stack([np.r_[0, np.cumsum(dx)], np.r_[0, np.cumsum(dy)], np.zeros(P)])
    delta[:, :2] = signal.detrend(delta[:, :2], axis=0, type="linear")
    delta[:, :2] = signal.savgol_filter(delta[:, :2], min(501, P // 2 * 2 - 1), 3, axis=0)
    maxnorm = float(np.linalg.norm(delta[:, :2], axis=1).max())
    if maxnorm > 1.0:
        delta *= 1. / maxnorm
    stats = dict(targets=len(points), cutoff_hz=cutoff,
                 median_gradient_coherence=float(np.median(coherence)),
                 delta_rms_m=np.sqrt(np.mean(delta ** 2, axis=0)).tolist(),
                 delta_max_m=np.max(np.abs(delta), axis=0).tolist())
    return delta, stats


def estimate_update(data, pos, points, r_start, dr, k, prf, cutoff, *, mode=None):
    """fast_bp/autofocus.py:37-79 with the same signature: per-(target, pulse)
    contributions on the GPU (``bpmi_target_contributions``, FP64 geometry
    and phase by default, complex64 out as the reference stores them), then
    ``phase_gradient_update`` on the host.  Returns (delta [P, 3], stats)."""
    from .ffbp import target_contributions
    xi = target_contributions(data, pos, points, r_start, dr, k, mode=mode)
    return phase_gradient_update(xi, pos, points, k, prf, cutoff)
