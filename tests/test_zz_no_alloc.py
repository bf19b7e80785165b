"""The allocation-free event window (dxg2_events_window_noalloc) against the workspace
form (dxg2_events_window), through the C ABI: two calls in a row, the second on a
smaller input, each bit-identical to the workspace form.  Marked gpu."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _window(lib, noalloc, o, t, d, hw, t0, t1, torch, _lib):
    cap = max(int(t.shape[0]), 1)
    off = torch.full((hw + 1,), -1, dtype=torch.int32, device="cuda")
    times = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    data = torch.zeros(cap, dtype=d.dtype, device="cuda")
    code = _lib.DXG2_U16 if d.dtype == torch.uint16 else _lib.DXG2_U8
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = lib.dxg2_events_window_noalloc(_lib.ptr(o), _lib.ptr(t), _lib.ptr(d),
                                            int(t.shape[0]), hw, code, t0, t1, _lib.ptr(off),
                                            _lib.ptr(times), _lib.ptr(data), cap, st)
    else:
        scratch = torch.empty((hw + 1023) // 1024 + 1, dtype=torch.int32, device="cuda")
        rc = lib.dxg2_events_window(_lib.ptr(o), _lib.ptr(t), _lib.ptr(d), int(t.shape[0]), hw,
                                    code, t0, t1, _lib.ptr(scratch), _lib.ptr(off),
                                    _lib.ptr(times), _lib.ptr(data), cap, st)
    _lib.check(rc, "events window")
    torch.cuda.synchronize()
    m = int(off[hw].item())
    return off.cpu().numpy(), times[:m].cpu().numpy(), data[:m].cpu().numpy()


@pytest.mark.parametrize("dtype", ["uint8", "uint16"])
def test_noalloc_window_equals_workspace_form(dtype):
    import torch
    from dynamix_amd import _lib, synth
    lib = _lib.load()
    cases = [((700, 900), 3000, 0.01, 250, 2800),   # 630,000 pixels: 616 scan blocks
             ((40, 50), 500, 0.05, 17, 400)]        # then a smaller input, same array
    for shape, n, mean, t0, t1 in cases:
        print("[case] noalloc window %s shape=%s n=%d [%d,%d)" % (dtype, shape, n, t0, t1),
              flush=True)
        sys.stdout.flush()
        t, d, o = synth.speckle_events_device(shape, n, mean, 50.0, seed=n, dtype=np.dtype(dtype))
        hw = shape[0] * shape[1]
        ref = _window(lib, False, o, t, d, hw, t0, t1, torch, _lib)
        got = _window(lib, True, o, t, d, hw, t0, t1, torch, _lib)
        assert ref[0][-1] > 0
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b)
        # the window itself, from the host copy of the list
        tn, dn = t.cpu().numpy(), d.cpu().numpy()
        on = o.cpu().numpy().view(np.uint32).astype(np.int64)
        pix = np.repeat(np.arange(hw), np.diff(on))
        keep = (tn >= t0) & (tn < t1)
        np.testing.assert_array_equal(got[1], tn[keep] - t0)
        np.testing.assert_array_equal(got[2], dn[keep])
