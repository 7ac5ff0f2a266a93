"""Seeded random-ellipse phantoms and their exact parallel-beam projections.

Geometry is algotom's: a pixel (x, y) of an (n, n) image (row y downward)
projects at angle theta onto detector position
    t = (x - ic) cos(theta) + (ic - y) sin(theta),   ic = 0.5 (n - 1),
i.e. the column ``center + t`` of the sinogram, which is the convention of
the back-projector (reconstruction.py:170-172 with theta -> -theta) and of
make_elliptic_mask (algotom/util/simulation.py:40-79, y up, angle in
degrees).  The projection of an ellipse is its textbook closed form, so
these sinograms need no forward projector.  Computation uses torch on any
device (float64).
"""

import numpy as np


def random_ellipses(size, count=20, seed=0, axis_range=(5.0, 120.0),
                    density_range=(0.1, 1.0)):
    """(count, 6) float64 rows (x0, y0, a, b, phi_deg, rho): centres uniform
    in a disk of radius 0.35*size, full axes U[axis_range] * size/512,
    rotation U[0, 180) degrees, density U[density_range]."""
    rng = np.random.default_rng(seed)
    r = 0.35 * size * np.sqrt(rng.uniform(0.0, 1.0, count))
    t = rng.uniform(0.0, 2.0 * np.pi, count)
    scale = size / 512.0
    axes = rng.uniform(axis_range[0], axis_range[1], (count, 2)) * scale
    phi = rng.uniform(0.0, 180.0, count)
    rho = rng.uniform(density_range[0], density_range[1], count)
    return np.column_stack([r * np.cos(t), r * np.sin(t), 0.5 * axes[:, 0],
                            0.5 * axes[:, 1], phi, rho])


def ellipse_image(params, size):
    """The phantom sampled at pixel centres, (size, size) float32."""
    ic = 0.5 * (size - 1)
    xs = np.arange(size) - ic
    X, Y = np.meshgrid(xs, -xs)
    img = np.zeros((size, size))
    for x0, y0, a, b, phi, rho in np.asarray(params):
        p = np.deg2rad(phi)
        u = (X - x0) * np.cos(p) + (Y - y0) * np.sin(p)
        v = -(X - x0) * np.sin(p) + (Y - y0) * np.cos(p)
        img += rho * ((u / a) ** 2 + (v / b) ** 2 <= 1.0)
    return img.astype(np.float32)


def ellipse_image_torch(params, size, device="cpu"):
    """ellipse_image as a torch float32 tensor computed on ``device`` (the
    same float64 expressions, so the same mask)."""
    import torch
    ic = 0.5 * (size - 1)
    xs = torch.arange(size, dtype=torch.float64, device=device) - ic
    X = xs[None, :]
    Y = -xs[:, None]
    img = torch.zeros((size, size), dtype=torch.float64, device=device)
    for x0, y0, a, b, phi, rho in np.asarray(params):
        p = np.deg2rad(phi)
        u = (X - x0) * np.cos(p) + (Y - y0) * np.sin(p)
        v = -(X - x0) * np.sin(p) + (Y - y0) * np.cos(p)
        inside = ((u / a) ** 2 + (v / b) ** 2 <= 1.0).to(torch.float64)
        img += inside * float(rho)
    return img.to(torch.float32)


def ellipse_sinogram(params, angles, ncol, center, device="cpu"):
    """Exact line integrals, (len(angles), ncol) torch float64 on device.

    Column k samples t = k - center."""
    import torch
    prm = torch.as_tensor(np.asarray(params), dtype=torch.float64,
                          device=device)
    th = torch.as_tensor(np.asarray(angles, dtype=np.float64),
                         device=device)[:, None]
    t = (torch.arange(ncol, dtype=torch.float64, device=device)
         - float(center))[None, :]
    out = torch.zeros((th.shape[0], ncol), dtype=torch.float64, device=device)
    for x0, y0, a, b, phi, rho in prm:
        p = torch.deg2rad(phi)
        s2 = (a * torch.cos(th - p)) ** 2 + (b * torch.sin(th - p)) ** 2
        t0 = x0 * torch.cos(th) + y0 * torch.sin(th)
        q = s2 - (t - t0) ** 2
        out += torch.where(q > 0, 2.0 * rho * a * b * torch.sqrt(
            torch.clamp(q, min=0.0)) / s2, torch.zeros_like(q))
    return out


def transmission(proj, photons=1.0e4, max_attenuation=3.0, seed=0):
    """Noisy transmission Poisson(photons * exp(-k p)) / photons with k such
    that max(k p) = max_attenuation; torch in, float32 torch out."""
    import torch
    k = max_attenuation / float(proj.max())
    gen = torch.Generator(device=proj.device)
    gen.manual_seed(int(seed))
    lam = photons * torch.exp(-k * proj)
    return (torch.poisson(lam, generator=gen) / photons).to(torch.float32)
