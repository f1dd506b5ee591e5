"""hifr_norm_residual_noalloc (the allocation-free residual norms) against the
allocating entry point hifr_norm_residual on the same residuals: two calls in a
row, each smaller than the one before, sharing the library's device scratch array;
plus the arguments it rejects.  Device inputs and outputs are torch tensors (torch
only provides the memory and the stream)."""
import ctypes as C

import numpy as np
import pytest

import cases
import hifr

pytestmark = pytest.mark.gpu

# (fixture, metric storage): 27 hexes P=3, then 12 hexes P=2, then 16 quads P=3
CASES = [("hex_p3_ns_perturbed_3x3x3", 1), ("hex_p2_euler_3x2x2", 0), ("quad_p3_euler_vortex_4x4", 0)]


def _lib():
    L = hifr.lib()
    L.hifr_norm_residual_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_int, C.c_int, C.c_int,
                                             C.c_double, C.c_void_p, C.c_void_p]
    return L


def _device_inputs(name, metrics):
    import torch
    g = cases.load_golden(name)
    s = cases.solver_of(g, hifr_metrics=metrics)
    s.set_state(g["u0"])
    div = s.residual()
    per_elem = s.metrics_per_element()
    detj = s.device_metrics()["detjac_upts"]
    detj = detj[0, :].copy() if per_elem else detj.reshape(-1, order="F").copy()
    d_div = torch.from_numpy(np.ascontiguousarray(div.reshape(-1, order="F"))).cuda()
    d_detj = torch.from_numpy(np.ascontiguousarray(detj)).cuda()
    return s, d_div, d_detj, (s.nu if per_elem else 1)


@pytest.mark.parametrize("norm_type", [0, 1, 2])
def test_noalloc_norms_equal_allocating_entry_point(require_gpu, norm_type):
    import torch
    L = _lib()
    stream = torch.cuda.current_stream().cuda_stream
    for name, metrics in CASES:
        print("case", name, "norm_type", norm_type, flush=True)
        s, d_div, d_detj, ppd = _device_inputs(name, metrics)
        want = s.norms(norm_type)
        nf = s.nf
        d_out = torch.full((nf,), float("nan"), dtype=torch.float64, device="cuda")
        rc = L.hifr_norm_residual_noalloc(d_div.data_ptr(), d_detj.data_ptr(), s.nu * s.n_eles, ppd, nf, norm_type,
                                          0.0, stream, d_out.data_ptr())
        assert rc == 0
        torch.cuda.synchronize()
        got = d_out.cpu().numpy()
        # the same partial sums, combined in the same order (max and L1 exactly)
        if norm_type < 2:
            assert np.array_equal(got, want), (got, want)
        else:
            np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)


def test_noalloc_rejects_what_the_scratch_cannot_hold(require_gpu):
    import torch
    L = _lib()
    d = torch.zeros(64, dtype=torch.float64, device="cuda")
    p = d.data_ptr()
    assert L.hifr_norm_residual_noalloc(p, p, 8, 1, 3, 2, 0.0, None, p) == 1
    assert L.hifr_norm_residual_noalloc(p, p, 8, 1, 6, 2, 0.0, None, p) == 1
    assert L.hifr_norm_residual_noalloc(p, p, 0, 1, 5, 2, 0.0, None, p) == 1
    assert L.hifr_norm_residual_noalloc(p, p, 8, 1, 5, 3, 0.0, None, p) == 1
    assert L.hifr_norm_residual_noalloc(None, p, 8, 1, 5, 2, 0.0, None, p) == 1
