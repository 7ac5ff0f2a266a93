"""Differentiable FBP: torch.autograd.Function with a hand-written adjoint.

Forward (per slice, reconstruction.py:362-493 without the log):
    y = sigma * mask * B_c(F s),   sigma = float32(pi) / float32(nrow - 1)
with F the windowed ramp filter (ath_ramp_filter) and B_c the
linear-interpolation back-projection at rotation centre c
(ath_back_projection, float64 sums).

Backward:
    dL/ds = F^T B_c^T (sigma * mask * dL/dy)
        B_c^T: ath_forward_projection (the back-projector's own weights,
               deterministic gather)
        F^T:   ath_ramp_filter_adjoint (correlation with the taps, transpose
               of the pad index map)
    dL/dc = sum(dL/dy * sigma * mask * sum_i (Fs)_i[d+1] - (Fs)_i[d]),
            d = floor f: ath_back_projection in mode ATH_ACC_DERIV (the
            interpolation is piecewise linear in f, and df/dc = 1).

Checkpointing: ``checkpoint="filtered"`` keeps the filtered sinogram from the
forward pass (needed by dL/dc); ``checkpoint="input"`` keeps only the input
and re-filters in backward, trading one filter launch for the memory of a
second copy of the stack; ``checkpoint="derivative"`` computes the centre
derivative image in the forward pass, in the same launch as the
back-projection (ath_back_projection_with_derivative: the positions and
staged samples are shared, each output bit-identical to its separate pass),
and keeps that image instead, so backward runs no derivative pass.
"""

import numpy as np

from algotom_amd import _native as nat
from algotom_amd.rec import reconstruction as rec


class FBPGeometry:
    """Fixed part of a differentiable FBP: angles, pad, window, mask."""

    def __init__(self, nrow, ncol, angles=None, ratio=1.0, filter_name="hann",
                 pad=None, pad_mode="edge", checkpoint="filtered",
                 center_for_ratio=None):
        import torch
        if pad_mode not in nat.PAD_MODES:
            raise NotImplementedError("pad_mode %r" % pad_mode)
        if checkpoint not in ("filtered", "input", "derivative"):
            raise ValueError("checkpoint must be 'filtered', 'input' or "
                             "'derivative'")
        self.dev = nat.device()
        self.lib = nat.load()
        self.nrow, self.ncol = int(nrow), int(ncol)
        if angles is None:
            angles = np.deg2rad(np.linspace(0.0, 180.0, nrow))
        if len(angles) != nrow:
            raise ValueError("!!!Number of angles is not the same as the row "
                             "number of the sinogram!!!")
        self.angles = rec._angles_device(angles, self.dev)
        self.pad = min(int(0.15 * ncol), 150) if pad is None else int(pad)
        self.pad_mode = pad_mode
        self.checkpoint = checkpoint
        width = ncol + 2 * self.pad
        taps = rec._window_taps(None, nrow, width, filter_name)
        self.fft_len = nat.check(self.lib.ath_filter_fft_length(ncol, self.pad),
                                 "ath_filter_fft_length")
        self.taps = taps
        # reversed taps g[(W - n) mod W] for the correlation of the adjoint
        rev = np.ascontiguousarray(np.roll(taps[:, ::-1], 1, axis=1))
        self.fft_len_adj = 64
        while self.fft_len_adj < 2 * width:
            self.fft_len_adj *= 2
        self.plan_adj, self.nwin_adj, _ = rec._filter_plan(rev, self.dev,
                                                           self.fft_len_adj)
        self.mask_radius = -1.0
        if ratio is not None:
            r = ratio
            if r == 0.0:
                if center_for_ratio is None:
                    raise ValueError("ratio=0 needs center_for_ratio")
                c = center_for_ratio
                r = min(c, ncol - c) / (0.5 * ncol)
            self.mask_radius = float(r * (ncol // 2))
        self.sigma = float(np.float32(np.pi)) / float(np.float32(nrow - 1))
        self._dtype = torch.float32

    # ---- kernels --------------------------------------------------------
    def filter(self, sino):
        return rec._filter_stack(sino, self.pad, self.pad_mode, False,
                                 self.taps, self._dtype)

    def backproject(self, filt, center, acc=nat.ACC_F64):
        return rec._backproject_stack(filt, self.angles, center, acc,
                                      mask_radius=self.mask_radius,
                                      apply_scale=True)

    def backproject_with_derivative(self, filt, center):
        """(sigma * mask * B_c filt, sigma * mask * dB_c/dc filt) from one
        launch."""
        import torch
        nrow, nslice, ncol = filt.shape
        out = torch.empty((ncol, nslice, ncol), dtype=torch.float32,
                          device=self.dev)
        der = torch.empty_like(out)
        nat.check(self.lib.ath_back_projection_with_derivative(
            nat.ptr(filt), nrow, nslice, ncol, nat.ptr(self.angles),
            float(np.float32(center)), 0, self.mask_radius, 1, nat.ptr(out),
            nat.ptr(der), nat.stream_ptr(self.dev)),
            "ath_back_projection_with_derivative")
        return out, der

    def forward_project(self, image, center):
        """sigma * B^T(mask * image): (ncol, nslice, ncol) -> (nrow, ...)."""
        import torch
        ncol, nslice, _ = image.shape
        out = torch.empty((self.nrow, nslice, ncol), dtype=torch.float32,
                          device=self.dev)
        # a transposed copy of the image lets every angle stage contiguous
        # rows (same bits as ath_forward_projection)
        nbytes = self.lib.ath_forward_projection_workspace_bytes(nslice, ncol)
        work = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
        nat.check(self.lib.ath_forward_projection_ws(
            nat.ptr(image), self.nrow, nslice, ncol, nat.ptr(self.angles),
            float(np.float32(center)), self.mask_radius, self.sigma,
            nat.ptr(work), nbytes, nat.ptr(out), nat.stream_ptr(self.dev)),
            "ath_forward_projection_ws")
        return out

    def filter_adjoint(self, u):
        import torch
        nrow, nslice, ncol = u.shape
        out = torch.empty_like(u)
        nat.check(self.lib.ath_ramp_filter_adjoint(
            nat.ptr(u), nrow, nslice, ncol, self.pad,
            nat.PAD_MODES[self.pad_mode], nat.ptr(self.plan_adj),
            self.fft_len_adj, self.nwin_adj, nat.ptr(out),
            nat.stream_ptr(self.dev)), "ath_ramp_filter_adjoint")
        return out


def _function():
    import torch

    class FBPFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, sino, center, geom):
            s = sino.detach().to(torch.float32).contiguous()
            c = float(center.detach().double().item())
            filt = geom.filter(s)
            ctx.geom = geom
            ctx.center = c
            if geom.checkpoint == "derivative":
                if ctx.needs_input_grad[1]:
                    out, deriv = geom.backproject_with_derivative(filt, c)
                    ctx.save_for_backward(deriv)
                else:
                    out = geom.backproject(filt, c)
                return out
            out = geom.backproject(filt, c)
            ctx.save_for_backward(filt if geom.checkpoint == "filtered"
                                  else s)
            return out

        @staticmethod
        def backward(ctx, grad):
            geom = ctx.geom
            saved = ctx.saved_tensors[0] if ctx.saved_tensors else None
            g = grad.detach().to(torch.float32).contiguous()
            g_sino = g_center = None
            if ctx.needs_input_grad[0]:
                g_sino = geom.filter_adjoint(geom.forward_project(g,
                                                                 ctx.center))
            if ctx.needs_input_grad[1]:
                if geom.checkpoint == "derivative":
                    deriv = saved
                else:
                    filt = saved if geom.checkpoint == "filtered" else \
                        geom.filter(saved)
                    deriv = geom.backproject(filt, ctx.center, nat.ACC_DERIV)
                g_center = torch.sum(g.double() * deriv.double()).reshape(())
            return g_sino, g_center, None

    return FBPFunction


_FBP = None


def fbp_differentiable(sinogram, center, geom):
    """y = FBP(sinogram; center) with gradients for both arguments.

    sinogram: (nrow, ncol) or (nrow, nslice, ncol) float32 tensor on the
    GPU; center: 0-d tensor (requires_grad for dL/dcenter).  Returns
    (ncol, ncol) or (ncol, nslice, ncol).
    """
    import torch
    global _FBP
    if _FBP is None:
        _FBP = _function()
    two_d = sinogram.dim() == 2
    s = sinogram.unsqueeze(1) if two_d else sinogram
    if not torch.is_tensor(center):
        center = torch.tensor(float(center), dtype=torch.float64)
    out = _FBP.apply(s, center, geom)
    return out[:, 0, :] if two_d else out


def total_variation(img, eps=1e-8):
    """Isotropic TV of (ncol, nslice, ncol) or (ncol, ncol) images."""
    import torch
    x = img if img.dim() == 3 else img.unsqueeze(1)
    dy = x[1:, :, :-1] - x[:-1, :, :-1]
    dx = x[:-1, :, 1:] - x[:-1, :, :-1]
    return torch.sum(torch.sqrt(dx * dx + dy * dy + eps))
