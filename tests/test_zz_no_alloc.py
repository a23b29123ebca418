"""msm_grid2grid_device_noalloc: grid2grid on device pointers with its
scratch in the library's static device array (no allocation).

Compared with the exported grid2grid (host arrays, per-call device buffers)
on the same padded level grid and stencil: the same kernels on the same
inputs, so the results agree bit for bit.  Two calls in a row, the second
on a smaller grid (the array keeps the first call's contents); the second
call leaves the free device memory unchanged; a grid beyond the array's
limit is refused before anything is launched.
"""
import ctypes

import numpy as np
import pytest

from helpers import load_dataset
from test_gpu_boundary import _reference_padded

pytestmark = pytest.mark.gpu


class _Triple(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int)]


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_noalloc_grid2grid_matches_grid2grid(torch_cuda):
    torch = torch_cuda
    from msmpb_amd import MSM, _lib
    lib = _lib.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    lib.grid2grid.argtypes = [_lib.FFP, ctypes.c_int, _Triple, dp, dp, _Triple, dp]
    lib.grid2grid.restype = None
    lib.msm_grid2grid_device_noalloc.argtypes = [_lib.FFP, ctypes.c_int, _Triple, ctypes.c_void_p,
                                                 ctypes.c_void_p, _Triple, ctypes.c_void_p,
                                                 ctypes.c_void_p]
    lib.msm_grid2grid_device_noalloc.restype = ctypes.c_int
    edges, q, pos, _ = load_dataset("spceN300")
    print("[no-alloc] grid2grid on device pointers, spceN300 L=2 nu=6", flush=True)
    m = MSM(pos, q, edges, nu=6, levels=2, top_grid=6, cutoff=9.0)
    try:
        nu, dmax = m.nu, 2 * m.nlim + 1
        rng = np.random.default_rng(17)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        # first the FF's own level 1, then a smaller grid with a stencil of its own
        cases = [((12, 12, 12), m.stencil(1)), ((4, 4, 5), rng.normal(size=(4, 4, 5)))]
        for call, (g, kh) in enumerate(cases):
            g = np.array(g)
            sd = np.minimum(dmax, g)
            qpad = np.ascontiguousarray(_reference_padded(rng.normal(size=tuple(g)), nu, dmax))
            khc = np.ascontiguousarray(kh, dtype=np.float64)
            el0 = rng.normal(size=tuple(g))
            T, S = _Triple(*[int(v) for v in g]), _Triple(*[int(v) for v in sd])
            el_h = el0.copy()
            lib.grid2grid(m._ff, 1, T, el_h.ctypes.data_as(dp), qpad.ctypes.data_as(dp), S,
                          khc.ctypes.data_as(dp))
            d_el = torch.tensor(el0, device="cuda")
            d_ql = torch.tensor(qpad, device="cuda")
            d_kh = torch.tensor(khc, device="cuda")
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info(0)[0]
            rc = lib.msm_grid2grid_device_noalloc(m._ff, 1, T, d_el.data_ptr(), d_ql.data_ptr(), S,
                                                  d_kh.data_ptr(), stream)
            assert rc == 0, _lib.last_error()
            torch.cuda.synchronize()
            if call == 1:  # the scratch array's address was looked up by the first call
                assert torch.cuda.mem_get_info(0)[0] == free0
            assert np.array_equal(d_el.cpu().numpy(), el_h), tuple(g)
            assert not np.array_equal(el_h, el0)
        # beyond the static array: refused on the host, nothing launched
        big = _Triple(81, 81, 81)
        rc = lib.msm_grid2grid_device_noalloc(m._ff, 1, big, d_el.data_ptr(), d_ql.data_ptr(),
                                              _Triple(4, 4, 4), d_kh.data_ptr(), stream)
        assert rc == _lib.MSM_ERR_ARG and "exceed" in _lib.last_error()
    finally:
        m.close()
