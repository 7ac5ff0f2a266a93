"""NumPy models of the HIP kernels' index arithmetic (CPU, test-only).

These mirror algotom_amd/csrc/ramp_filter.hip and backproject.hip step by
step (same pass structure, same index formulas, same table indices) so that
the kernels' addressing can be checked on the CPU against the oracle before
anything runs on a GPU.  They are not the oracle: the oracle is oracle/.
"""

import math

import numpy as np


def twiddles(L):
    n = np.arange(L)
    return np.exp(-2j * np.pi * n / L)


def lds_fft(buf, tw, direction):
    """Model of lds_fft<DIR>: radix-4 Stockham passes then one radix-2."""
    buf = buf.copy()
    M = buf.size
    L = 2 * M
    ns = 1
    while ns * 4 <= M:
        nb = M >> 2
        tstep = L // (ns * 4)
        j = np.arange(nb)
        v = np.stack([buf[j + r * nb] for r in range(4)])
        k = j & (ns - 1)
        for r in range(1, 4):
            w = tw[r * k * tstep]
            if direction > 0:
                w = np.conj(w)
            v[r] = v[r] * w
        a0 = v[0] + v[2]
        a1 = v[0] - v[2]
        a2 = v[1] + v[3]
        d = v[1] - v[3]
        a3 = -1j * d if direction < 0 else 1j * d
        out = [a0 + a2, a1 + a3, a0 - a2, a1 - a3]
        base = (j - k) * 4 + k
        for r in range(4):
            buf[base + r * ns] = out[r]
        ns *= 4
    if ns < M:
        nb = M >> 1
        j = np.arange(nb)
        x0 = buf[j].copy()
        x1 = buf[j + nb].copy()
        w = tw[2 * j]
        if direction > 0:
            w = np.conj(w)
        x1 = x1 * w
        buf[j] = x0 + x1
        buf[j + ns] = x0 - x1
    return buf


def split_real(zk, zm, wk):
    e = 0.5 * (zk + np.conj(zm))
    o = -1j * (0.5 * (zk - np.conj(zm)))
    wo = wk * o
    return e + wo, np.conj(e - wo)


def kernel_spectrum(g, L):
    """Model of kernel_spectrum_kernel: rfft_L([g, g, 0...]) / M."""
    M = L // 2
    W = g.size
    tw = twiddles(L)
    x = np.zeros(L)
    n = np.arange(min(2 * W, L))
    x[n] = g[np.where(n < W, n, n - W)]
    buf = x[0::2] + 1j * x[1::2]
    z = lds_fft(buf, tw, -1)
    G = np.zeros(M + 1, dtype=complex)
    k = np.arange(M // 2 + 1)
    km = (M - k) & (M - 1)
    xk, xm = split_real(z[k], z[km], tw[k])
    G[k] = xk / M
    G[M - k] = xm / M
    return G


def ramp_filter_row(p_row, g, pad, L):
    """Model of ramp_filter_kernel on one already padded row (length W)."""
    M = L // 2
    W = p_row.size
    ncol = W - 2 * pad
    tw = twiddles(L)
    G = kernel_spectrum(g, L)
    x = np.zeros(L)
    x[:W] = p_row
    z = lds_fft(x[0::2] + 1j * x[1::2], tw, -1)
    out = z.copy()
    for k in range(M // 2 + 1):
        km = (M - k) & (M - 1)
        wk = tw[k]
        xk, xm = split_real(z[k], z[km], wk)
        yk = xk * G[k]
        ym = xm * G[M - k]
        ek = 0.5 * (yk + np.conj(ym))
        ok = 1j * (np.conj(wk) * (0.5 * (yk - np.conj(ym))))
        out[k] = ek + ok
        if k != 0 and k != km:
            em = 0.5 * (ym + np.conj(yk))
            om = -1j * (wk * (0.5 * (ym - np.conj(yk))))
            out[km] = em + om
    zy = lds_fft(out, tw, 1)
    y = np.empty(L)
    y[0::2] = zy.real
    y[1::2] = zy.imag
    off = W + pad
    return y[off:off + ncol]


def fft_length(ncol, pad):
    need = 2 * (ncol + 2 * pad) - pad
    L = 64
    while L < need:
        L *= 2
    return L


def bp_span(tx, ty):
    return int(math.ceil(math.sqrt((tx - 1) ** 2 + (ty - 1) ** 2) - 1e-12)) + 3


def bp_tile_footprints(ncol, angles, center, py=8, tx=64):
    """Model of bp_tiled_kernel's per-(tile, angle) staging bounds.

    Returns the largest staged index any pixel of any tile reads
    (d - dlo on the range-free path, for pixels the kernel accumulates) and
    the span the kernel allocates.
    """
    ty = 4 * py
    span = bp_span(tx, ty)
    ic = 0.5 * (ncol - 1)
    th = -np.asarray(angles, dtype=np.float32).astype(np.float64)
    cs = np.cos(th).astype(np.float32).astype(np.float64)
    sn = np.sin(th).astype(np.float32).astype(np.float64)
    worst = 0
    for y0 in range(0, ncol, ty):
        for x0 in range(0, ncol, tx):
            xs = np.arange(x0, x0 + tx) - ic
            ys = np.arange(y0, y0 + ty) - ic
            cx = np.array([x0, x0 + tx - 1]) - ic
            cy = np.array([y0, y0 + ty - 1]) - ic
            for c, s in zip(cs, sn):
                corners = (cy[:, None] * s + (cx[None, :] * c)) + center
                fmn, fmx = corners.min(), corners.max()
                dlo = math.floor(fmn)
                f = (ys[:, None] * s + (xs[None, :] * c)) + center
                assert f.min() >= fmn and f.max() <= fmx
                inner = fmn >= 0.0 and fmx <= ncol - 1
                if inner:
                    d = np.trunc(f).astype(np.int64)
                    worst = max(worst, int((d - dlo).max()))
                    assert int((d - dlo).min()) >= 0
    return worst, span


def bp_rec_planes(rec_bytes, sb, planes):
    """RecPlanes<Rec, SB, PLANES>: (P, NP) = slices per plane, planes."""
    per16 = 16 // rec_bytes
    p = per16 if planes and 1 <= per16 < sb else sb
    return p, sb // p


def bp_rec_stage_index(al, k, js, span, sb, p):
    """Record index bp_tile's staging loop writes for (angle, column, slice)."""
    return al * span * sb + k * p + (js // p) * span * p + js % p


def bp_rec_read_index(al, d, dlo, js, span, sb, p):
    """Record index inner_angle reads: the wave-uniform offset (a multiple of
    P), plus d * P, plus the slice's offset within the column."""
    npl = sb // p
    off = (al * span * npl - dlo) * p
    return off + d * p + (js // p) * span * p + js % p


# ---- forward_project_hat_kernel's staging ---------------------------------
def fp_block(p, nrow, ncol, nslice, fsb):
    """forward_project_hat_kernel's workgroup id -> (bin block, angle block,
    slice block): every XCD (id % 8) gets a contiguous run of the order
    with the angle block fastest."""
    nab, nbb = (nrow + 3) // 4, (ncol + 63) // 64
    nsb = (nslice + fsb - 1) // fsb
    total = nab * nbb * nsb
    x, slot = p & 7, p >> 3
    per, rem = total >> 3, total & 7
    l = x * per + min(x, rem) + slot
    return (l // nab) % nbb, l % nab, (l // nab) // nbb


def fp_chunk_is_empty(ncol, use_mask, mask_r2, mc, m0, m1, ka, kb):
    """The kernel's chunk_is_empty: no pixel of lines m0..m1 and columns
    ka..kb (plus slack) inside the image and the mask."""
    if kb < -2 or ka > ncol + 1:
        return True
    if not use_mask:
        return False
    dm = m0 - mc if mc < m0 else (mc - m1 if mc > m1 else 0)
    rest = mask_r2 - float(dm) * float(dm)
    if rest < 0.0:
        return True
    h = math.sqrt(rest) + 2.0
    return kb < mc - h or ka > mc + h


# ---- summation order of metrics.hip (numpy's float32 mean) ---------------
def np_pairwise_leaf(a, f=np.float32):
    """np_pairwise_leaf: <= 128 values, 8 partial sums of type f."""
    n = len(a)
    if n < 8:
        r = f(0)
        for x in a:
            r = f(r + x)
        return r
    r = [f(a[j]) for j in range(8)]
    i = 8
    while i < n - (n % 8):
        for j in range(8):
            r[j] = f(r[j] + a[i + j])
        i += 8
    res = f(f(f(r[0] + r[1]) + f(r[2] + r[3]))
            + f(f(r[4] + r[5]) + f(r[6] + r[7])))
    while i < n:
        res = f(res + a[i])
        i += 1
    return res


def np_pairwise(a, f=np.float32):
    """np_pairwise_serial: the recursive split at n/2 rounded to 8."""
    n = len(a)
    if n <= 128:
        return np_pairwise_leaf(a, f)
    n2 = n // 2
    n2 -= n2 % 8
    return f(np_pairwise(a[:n2], f) + np_pairwise(a[n2:], f))


def np_mean_f32(x, chunk=8192):
    """sharpness_kernel's reduction: sequential float32 sum of the pairwise
    sums of consecutive `chunk`-value pieces; full pieces as 64 leaves of
    128 combined by a perfect tree (vectorised here), the tail recursively;
    mean = float32(float64(sum) / n)."""
    flat = np.ascontiguousarray(x, dtype=np.float32).ravel()
    n = flat.size
    nfull = n // chunk
    sums = []
    if nfull:
        blk = flat[:nfull * chunk].reshape(nfull, 64, 16, 8)
        r = blk[:, :, 0, :].copy()
        for i in range(1, 16):
            r = (r + blk[:, :, i, :]).astype(np.float32)
        leaf = ((r[..., 0] + r[..., 1]) + (r[..., 2] + r[..., 3])) + \
            ((r[..., 4] + r[..., 5]) + (r[..., 6] + r[..., 7]))
        while leaf.shape[1] > 1:
            leaf = (leaf[:, 0::2] + leaf[:, 1::2]).astype(np.float32)
        sums.extend(leaf[:, 0].tolist())
    if n % chunk:
        sums.append(np_pairwise(flat[nfull * chunk:]))
    total = np.float32(0)
    for s in sums:
        total = np.float32(total + np.float32(s))
    return np.float32(np.float64(total) / np.float64(n))
