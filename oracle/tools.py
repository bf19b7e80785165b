"""Loop restatements of the reference's host tools (TEST INFRASTRUCTURE ONLY -- see
oracle/__init__.py): element by element, in the order dynamix/tools/tools.py does it,
to pin dynamix_amd/tools.py.

radi_loops      tools.py:14-50   radial average on integer radii + interpolated model
make_q_loops    tools.py:262-295 per-pixel q and annulus test
cftomt_loops    tools.py:146-172 multitau resampling, one row at a time

Parity pinning: radi, cftomt and format_result are also pinned directly by golden
vectors written by running those reference functions (tests/golden/make_golden.py;
tests/test_tools.py compares dynamix_amd/tools.py with them).  make_q needs its ini-file
configuration and is pinned by make_q_loops and the closed forms of tests/test_tools.py
only.
"""
import math

import numpy as np


def radi_loops(saxs, mask, cx, cy):
    h, w = saxs.shape
    q = np.zeros((h, w), np.float32)
    qh = np.zeros((h, w), np.int64)
    for y in range(h):
        for x in range(w):
            q[y, x] = np.float32(math.sqrt((x - cx) ** 2 + (y - cy) ** 2))
            qh[y, x] = int(np.float32(q[y, x] + np.float32(0.5)))
    pts = [(qh[y, x], float(saxs[y, x])) for y in range(h) for x in range(w) if mask[y, x] < 1]
    lo = min(p[0] for p in pts)
    hi = max(p[0] for p in pts)
    nbin = hi - lo
    s = [0.0] * nbin
    c = [0] * nbin
    for v, val in pts:
        k = min(v - lo, nbin - 1)
        s[k] += val
        c[k] += 1
    rad = np.array([[lo + k, s[k] / c[k] if c[k] else float("nan")] for k in range(nbin)])
    vals = list(rad[:, 1]) + list(rad[-2:, 1])
    model = np.zeros((h, w), np.float32)
    r = q.copy()
    for y in range(h):
        for x in range(w):
            if mask[y, x] > 0:
                r[y, x] = 0
                continue
            fl = int(np.uint16(q[y, x]))
            f1 = np.float32(q[y, x] - np.float32(fl))
            ind = (fl - lo) % 65536
            if ind > hi:
                ind = 0
            model[y, x] = vals[ind + 1] * f1 + vals[ind] * (1 - f1)
    return rad, r, model


def make_q_loops(shape, cx, cy, mask, first_q, width_q, step_q, number_q, pix_size, distance,
                 wavelength):
    h, w = shape
    qp = np.linspace(first_q, first_q + (number_q - 1) * step_q, number_q)
    out = np.zeros((h, w), np.uint16)
    for y in range(h):
        for x in range(w):
            if mask[y, x] > 0:
                continue
            r = np.float32(math.sqrt((x - cx) ** 2 + (y - cy) ** 2))
            t = np.arctan(np.float32(np.float32(pix_size * r) / np.float32(distance)))
            qv = np.float32(4 * np.pi / wavelength) * np.sin(np.float32(t / np.float32(2.0)))
            for n in range(number_q):
                if abs(float(qv) - qp[n]) <= width_q / 2:
                    out[y, x] = n + 1
    return out


def cftomt_loops(d, par=16):
    rows = [list(r[:3]) for r in np.asarray(d)]
    out = rows[:par]
    level = rows[par:]
    while len(level) >= par:
        pairs = [[np.float32((a + b) / 2) if isinstance(a, np.float32) else (a + b) / 2
                  for a, b in zip(level[k], level[k + 1])] for k in range(len(level) - 1)]
        out.extend(pairs[k] for k in range(0, par, 2))
        level = [pairs[k] for k in range(par, len(pairs) - 1, 2)]
    return np.array(out, dtype=np.asarray(d).dtype)
