"""olpt_find_peaks_static (no context, no allocation) against olpt_find_peaks,
through the C-ABI: two calls in a row on the same GPU work array, the second
on a smaller image, each bit-identical to the allocating entry point; an image
beyond the array's maximum is refused before any GPU work."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _static_peaks(L, img: torch.Tensor, code: int, params, cap: int):
    from pyopenlpt_amd.context import ptr
    H, W = img.shape
    xy = torch.full((cap, 2), -1.0, dtype=torch.float64, device=img.device)
    n = torch.full((1,), -1, dtype=torch.int32, device=img.device)
    st = torch.cuda.current_stream(img.device)
    rc = L.olpt_find_peaks_static(ctypes.c_void_p(st.cuda_stream), ptr(img), code, H, W,
                                  ctypes.byref(params), ptr(xy), cap, ptr(n))
    assert rc == 0, L.olpt_last_error()
    st.synchronize()
    total = int(n.item())
    return xy[: min(total, cap)].cpu().numpy(), total


def test_static_peaks_equal_allocating_entry_point(gpu_available):
    from pyopenlpt_amd import _lib
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.context import Context, to_device_image
    L = _lib.lib()
    big = S.workload_c2()                                  # 1024^2 u16
    small = S.workload_c1(seed=11, n=600, size=256)        # 256^2 u8
    for i, w in enumerate((big, small)):
        H, W = w.images.shape[1:]
        print(f"[no-alloc] call {i + 1}: {H}x{W} {w.images.dtype}", flush=True)
        d, code = to_device_image(w.images)
        img = d[1].contiguous()
        ctx = Context(w.cams, H, W)
        want = ctx.find_peaks(img, code, w.params, cap=1 << 18)
        ctx.close()
        got, total = _static_peaks(L, img, code, w.params, 1 << 18)
        assert total == len(want)
        assert np.array_equal(got, want)
        # a capacity below the count: the total is still reported and the
        # first `cap` peaks are the allocating entry point's first ones
        cap = max(1, len(want) // 3)
        got2, total2 = _static_peaks(L, img, code, w.params, cap)
        assert total2 == len(want)
        assert np.array_equal(got2, want[:cap])


def test_static_peaks_rejects_oversized_images(gpu_available):
    from pyopenlpt_amd import _lib
    from pyopenlpt_amd import synth as S
    L = _lib.lib()
    p = S.default_params(0.04)
    n = torch.zeros(1, dtype=torch.int32, device="cuda")
    img = torch.zeros(8, dtype=torch.uint8, device="cuda")
    rc = L.olpt_find_peaks_static(None, ctypes.c_void_p(img.data_ptr()), _lib.OLPT_U8, 3, 4097,
                                  ctypes.byref(p), None, 0, ctypes.c_void_p(n.data_ptr()))
    assert rc == 102
    assert b"4096" in L.olpt_last_error()
