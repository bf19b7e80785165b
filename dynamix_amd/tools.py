"""The steps around the correlator: geometry -> qmask, and linear -> multitau results.

Host-side numpy (these are cheap next to the correlation and need no GPU):

radial_average  -- tools.radi (dynamix/tools/tools.py:14-50): azimuthal average of a 2-D
                   pattern about (cx, cy) on integer radii, and the pattern rebuilt from
                   it by linear interpolation in the radius.
q_map           -- the per-pixel momentum transfer of tools.make_q (tools.py:262-263):
                   q = 4 pi / lambda sin(arctan(pixel r / distance) / 2).
make_qmask      -- the qmask of tools.make_q (tools.py:276-295): number_q annuli of
                   width width_q centred on first_q + n step_q, later annuli winning where
                   they overlap, masked pixels 0.
multitau        -- tools.cftomt (tools.py:146-172): a linearly spaced correlation function
                   (columns lag, value, error) resampled onto multitau spacing with
                   `par` channels per level.
format_result   -- tools.format_result (tools.py:313-346): flat-field and speckle-contrast
                   corrections per q, multitau spacing, and the table for saving.
"""
import numpy as np


def radial_average(saxs, mask, cx, cy):
    """(rad, r, model): rad[:, 0] integer radii from the smallest unmasked one, rad[:, 1]
    the mean intensity of the unmasked pixels whose rounded radius equals it; r the
    float32 radius of every pixel (0 where masked); model the pattern rebuilt from rad by
    interpolating linearly in r (0 where masked)."""
    saxs = np.asarray(saxs)
    mask = np.asarray(mask)
    yy, xx = np.indices(saxs.shape)
    r = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2).astype(np.float32)
    ring = (r + 0.5).astype(np.int16)         # nearest integer radius
    keep = mask < 1
    r_in, v_in = ring[keep].astype(np.int64), saxs[keep]
    lo, hi = int(r_in.min()), int(r_in.max())
    edges = np.arange(lo, hi + 1)
    # histogram bins [e_k, e_k+1), the last one closed, as np.histogram counts them
    idx = np.minimum(r_in - lo, len(edges) - 2) if len(edges) > 1 else np.zeros_like(r_in)
    nbin = max(len(edges) - 1, 0)
    sums = np.bincount(idx, weights=v_in, minlength=nbin)[:nbin]
    counts = np.bincount(idx, minlength=nbin)[:nbin]
    rad = np.zeros((nbin, 2))
    rad[:, 0] = edges[:-1]
    with np.errstate(divide="ignore", invalid="ignore"):
        rad[:, 1] = sums / counts
    r = np.where(keep, r, np.float32(0))
    model = np.zeros_like(r)
    vals = np.concatenate([rad[:, 1], rad[-2:, 1]])
    # uint16 arithmetic, as tools.py:43 does it: a pixel whose floor(r) lies below the
    # first radius wraps to a large index, which the next rule sends to 0
    base = (r.astype(np.uint16).astype(np.int64) - int(edges[0])) % 65536
    base = np.where(keep, base, 0)
    base = np.where(base > edges[-1], 0, base)
    frac = r - r.astype(np.uint16)
    model[keep] = (vals[base[keep] + 1] * frac[keep] + vals[base[keep]] * (1 - frac[keep]))
    return rad, r, model


def q_map(r, pix_size, distance, wavelength):
    """Momentum transfer of a radius map in pixels (tools.py:262-263), in 1/[wavelength]."""
    theta2 = np.arctan(pix_size * np.asarray(r) / distance)
    return 4 * np.pi / wavelength * np.sin(theta2 / 2.0)


def make_qmask(shape, cx, cy, mask, first_q, width_q, step_q, number_q, pix_size, distance,
               wavelength):
    """qmask of tools.make_q (tools.py:276-295), same dtype (uint16): bin n (1-based) holds
    the pixels with |q - (first_q + (n-1) step_q)| <= width_q / 2, a later bin taking a
    pixel that two annuli share; masked pixels (mask > 0) are 0.  The radius map is the
    one radial_average builds (masked pixels at radius 0)."""
    mask = np.asarray(mask)
    yy, xx = np.indices(tuple(shape))
    r = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2).astype(np.float32)
    r[mask > 0] = 0
    q = q_map(r, pix_size, distance, wavelength)
    centres = np.linspace(first_q, first_q + (number_q - 1) * step_q, number_q)
    qmask = np.zeros(tuple(shape), np.uint16)
    for n, c in enumerate(centres, start=1):
        qmask[np.abs(q - c) <= width_q / 2] = n
    qmask[mask > 0] = 0
    return qmask


def multitau(cf, par=16):
    """Multitau spacing of a linearly spaced function (tools.cftomt, tools.py:146-172).

    cf: rows (lag, value, error).  The first `par` rows are kept; then, level by level,
    the remaining rows are averaged pairwise (rows k and k+1), every other average of
    the first `par` is kept (par / 2 points per level), and the averages from row `par`
    on, every other one and without the last, form the next level, until a level has
    fewer than `par` rows.  The arithmetic stays in cf's dtype, like the reference.
    """
    cf = np.asarray(cf)[:, :3]
    out = [cf[:par]]
    level = cf[par:]
    while level.shape[0] >= par:
        pairs = (level[:-1] + level[1:]) / 2
        out.append(pairs[0:par:2])
        level = pairs[par:-1:2]
    return np.concatenate(out, axis=0)


def format_result(result, qmask, flatfield, cdata, dt, ttcf_par):
    """tools.format_result (tools.py:313-346).

    result: CorrelationResult with .res (n_q, nframes), .dev and .trc.  For every q the
    correlation and its error are divided by the speckle correction
    <c^2>/<c>^2 * <f^2>/<f>^2 over the bin's pixels (cdata: smoothed pattern, f: flat
    field), put on lags dt, 2 dt, ... and resampled onto multitau spacing.  Returns
    (list of per-q (lag, g2, err) arrays, the table [lag, g2_1, err_1, g2_2, err_2, ...],
    trc).  As in the reference, the returned trc is the uncorrected result.trc.
    """
    nframes = result.res[0].size
    cf = np.zeros((nframes, 3), np.float32)
    cf[:, 0] = np.arange(1, nframes + 1, 1) * dt
    per_q, table = [], None
    for n, q in enumerate(range(1, int(np.max(qmask)) + 1)):
        sel = qmask == q
        ff, cd = flatfield[sel], cdata[sel]
        correction = (cd ** 2).mean() / cd.mean() ** 2 * ((ff ** 2).mean() / ff.mean() ** 2)
        cf[:, 1] = result.res[n, :] / correction
        cf[:, 2] = result.dev[n, :] / correction
        m = multitau(cf)
        per_q.append(m)
        table = m if table is None else np.append(table, m[:, 1:], axis=1)
    return per_q, table, result.trc
