"""Synthetic XPCS inputs: annular qmasks and AR(1) speckle photon-count stacks.

Speckle model: per pixel a complex Gaussian field E_t = rho E_{t-1} +
sqrt(1 - rho^2) w_t, w ~ CN(0, 1), rho = exp(-1/tau_c), E_0 ~ CN(0, 1); counts
I(t, p) = Poisson(mu_p |E_t|^2) clipped to the dtype.  g2(tau) of such data is
1 + beta exp(-2 tau / tau_c) for tau > 0 (Siegert relation, beta = 1 per pixel).

Small stacks are made on the host with numpy (PCG64, seeded); large stacks are made on
the GPU by dxg2_synth_speckle (Philox-4x32-10, seeded).  The two generators draw
different random streams; every check compares a correlator with the oracle on the
same stack, so the stacks need not agree.

Workloads (sizes of the benchmark configurations):
  small     256 x 256, 1000 frames, uint8, mean 0.05, tau_c 50, 10 equal annuli
  headline  514 x 1030, 10000 frames, uint8, mean 0.02, tau_c 200, 20 equal annuli
            out to the corner, six 8-row gap bands masked (48 rows, 9.3 %)
  uint16    2162 x 2068, 5000 frames, uint16, mean 0.5, 1 % hot pixels at mean 40,
            50 annuli of linearly growing width
  events    1030 x 1065, 100000 frames, mean 1e-3, 30 equal annuli, tau_c 200
"""
import ctypes

import numpy as np

WORKLOADS = {
    "small": dict(shape=(256, 256), nframes=1000, dtype="uint8", mean=0.05, tau_c=50.0,
                  n_bins=10, center=(128.0, 128.0), r_max=128.0, widths="equal", gaps=0),
    "headline": dict(shape=(514, 1030), nframes=10000, dtype="uint8", mean=0.02, tau_c=200.0,
                     n_bins=20, center=(257.0, 515.0), r_max=None, widths="equal", gaps=6),
    "uint16": dict(shape=(2162, 2068), nframes=5000, dtype="uint16", mean=0.5, tau_c=200.0,
                   n_bins=50, center=(1081.0, 1034.0), r_max=None, widths="linear", gaps=0,
                   hot_fraction=0.01, hot_mean=40.0),
    "events": dict(shape=(1030, 1065), nframes=100000, dtype="uint8", mean=1e-3, tau_c=200.0,
                   n_bins=30, center=(515.0, 532.0), r_max=None, widths="equal", gaps=0),
}


def annular_qmask(shape, n_bins, center=None, r_max=None, widths="equal", gaps=0):
    """Annuli about `center` (row, col); 0 = excluded.

    r_max None: out to the farthest corner (every pixel binned).  widths "equal": equal
    radial widths; "linear": width of annulus k proportional to k + 1.  gaps: number
    of evenly spaced 8-row horizontal bands set to 0 (detector module gaps).
    """
    h, w = shape
    cy, cx = center if center is not None else ((h - 1) / 2.0, (w - 1) / 2.0)
    yy, xx = np.mgrid[0:h, 0:w]
    r = np.hypot(yy - cy, xx - cx)
    if r_max is None:
        r_max = float(r.max()) * (1 + 1e-9)
    if widths == "equal":
        edges = np.linspace(0.0, r_max, n_bins + 1)
    elif widths == "linear":
        wk = np.arange(1, n_bins + 1, dtype=np.float64)
        edges = np.concatenate([[0.0], np.cumsum(wk) / wk.sum() * r_max])
    else:
        raise ValueError(widths)
    q = np.searchsorted(edges, r, side="right").astype(np.int32)
    q[(r >= r_max) | (q > n_bins)] = 0
    for k in range(gaps):
        r0 = int(round((k + 0.5) * h / gaps)) - 4
        q[max(r0, 0):max(r0, 0) + 8, :] = 0
    return q


def workload_qmask(name):
    c = WORKLOADS[name]
    return annular_qmask(c["shape"], c["n_bins"], c["center"], c["r_max"], c["widths"], c["gaps"])


def mean_map(shape, mean, hot_fraction=0.0, hot_mean=0.0, seed=0):
    """Per-pixel mean count per frame; a seeded `hot_fraction` of pixels at hot_mean."""
    m = np.full(shape, mean, dtype=np.float32)
    if hot_fraction > 0:
        rng = np.random.default_rng(seed ^ 0x5EED)
        hot = rng.random(shape) < hot_fraction
        m[hot] = hot_mean
    return m


def speckle_frames(shape, nframes, mean, tau_c, seed, dtype=np.uint8, mean_map_=None):
    """Host (numpy PCG64) AR(1) speckle stack, for small sizes."""
    rng = np.random.default_rng(seed)
    h, w = shape
    mu = np.full(shape, mean, np.float64) if mean_map_ is None else mean_map_.astype(np.float64)
    rho = np.exp(-1.0 / tau_c) if tau_c > 0 else 0.0
    c = np.sqrt(1.0 - rho * rho)
    s = np.sqrt(0.5)
    e = s * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    info = np.iinfo(dtype)
    out = np.empty((nframes, h, w), dtype=dtype)
    for t in range(nframes):
        if t:
            e = rho * e + c * s * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        out[t] = np.clip(rng.poisson(mu * (e.real ** 2 + e.imag ** 2)), 0, info.max)
    return out


def speckle_frames_device(shape, nframes, mean, tau_c, seed, dtype=np.uint8, mean_map_=None):
    """GPU AR(1) speckle stack as a torch tensor (dxg2_synth_speckle)."""
    import torch
    from . import _lib
    lib = _lib.load()
    dtype = np.dtype(dtype)
    tdt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.int8): torch.int8,
           np.dtype(np.uint16): torch.uint16}[dtype]
    mm = mean_map(shape, mean) if mean_map_ is None else np.asarray(mean_map_, np.float32)
    d_mean = torch.from_numpy(np.ascontiguousarray(mm, dtype=np.float32)).to("cuda")
    frames = torch.empty((nframes,) + tuple(shape), dtype=tdt, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.dxg2_synth_speckle(_lib.ptr(frames), _lib.DTYPE_CODES[dtype], nframes,
                                      shape[0], shape[1], _lib.ptr(d_mean), float(tau_c),
                                      int(seed) & 0xFFFFFFFFFFFFFFFF, stream),
               "dxg2_synth_speckle")
    return frames


def workload_frames_device(name, seed):
    """GPU frames of a named workload, with its hot-pixel population if it has one."""
    c = WORKLOADS[name]
    mm = mean_map(c["shape"], c["mean"], c.get("hot_fraction", 0.0), c.get("hot_mean", 0.0),
                  seed)
    return speckle_frames_device(c["shape"], c["nframes"], c["mean"], c["tau_c"], seed,
                                 np.dtype(c["dtype"]), mean_map_=mm)


def speckle_events_device(shape, nframes, mean, tau_c, seed, dtype=np.uint8, mean_map_=None):
    """GPU events of the same walk as speckle_frames_device, in the pixel-major layout:
    returns torch tensors (times int32, data, offsets uint32)."""
    import torch
    from . import _lib
    lib = _lib.load()
    dtype = np.dtype(dtype)
    tdt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint16): torch.uint16}[dtype]
    mm = mean_map(shape, mean) if mean_map_ is None else np.asarray(mean_map_, np.float32)
    d_mean = torch.from_numpy(np.ascontiguousarray(mm, dtype=np.float32)).to("cuda")
    hw = int(shape[0]) * int(shape[1])
    offsets = torch.empty(hw + 1, dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    nnz = ctypes.c_int64()
    args = (_lib.ptr(d_mean), shape[0], shape[1], nframes, float(tau_c),
            int(seed) & 0xFFFFFFFFFFFFFFFF, _lib.DTYPE_CODES[dtype], _lib.ptr(offsets))
    _lib.check(lib.dxg2_synth_events(*args, None, None, 0, ctypes.byref(nnz), stream),
               "dxg2_synth_events")
    n = nnz.value
    times = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    data = torch.empty(max(n, 1), dtype=tdt, device="cuda")
    _lib.check(lib.dxg2_synth_events(*args, _lib.ptr(times), _lib.ptr(data), max(n, 1),
                                     ctypes.byref(nnz), stream), "dxg2_synth_events")
    return times[:n], data[:n], offsets


def events_to_coo_device(times, data, offsets, shape, nframes):
    """Pixel-major events (torch, on the GPU) -> frame-sorted COO (pixel, frame, count)."""
    import torch
    from . import _lib
    lib = _lib.load()
    n = int(times.shape[0])
    pix = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    fr = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    cnt = torch.empty(max(n, 1), dtype=data.dtype, device="cuda")
    code = _lib.DXG2_U16 if data.dtype == torch.uint16 else _lib.DXG2_U8
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.dxg2_events_to_coo(_lib.ptr(times), _lib.ptr(data), _lib.ptr(offsets), n,
                                      shape[0], shape[1], nframes, code, _lib.ptr(pix),
                                      _lib.ptr(fr), _lib.ptr(cnt), stream), "dxg2_events_to_coo")
    return pix[:n], fr[:n], cnt[:n]
