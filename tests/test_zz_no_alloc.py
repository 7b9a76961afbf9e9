"""sop_chol_factor_noalloc: the blocked Cholesky on caller-owned device
memory, with its workspace in the library's static device array, equals the
allocating entry point sop_chol_factor bit for bit -- two calls in a row per
precision, the second on a smaller matrix, so the shared workspace keeps the
first call's contents when the second runs."""
import numpy as np
import pytest

import bench
import sopgpu as S

pytestmark = pytest.mark.gpu

MAX_ORDER = 12288


def _noalloc():
    return S.lib.sop_chol_factor_noalloc


def _factor_in_place(n, A, mode, ld_extra):
    """Factor A (n x n) in a device buffer of mp rows x (mp + ld_extra)
    floats whose padding rows and columns start as NaN."""
    import torch

    mp = (n + 63) // 64 * 64
    ld = mp + ld_extra
    buf = torch.full((mp, ld), float("nan"), dtype=torch.float32, device="cuda")
    buf[:n, :n] = torch.from_numpy(A).cuda()
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    print("case: no-alloc factor, order %d, ld %d, mode %d" % (n, ld, mode), flush=True)
    rc = _noalloc()(0, n, buf.data_ptr(), ld, mode, stream.cuda_stream)
    assert rc == 0, S.lib.sop_last_error()
    torch.cuda.synchronize()
    return np.tril(buf[:n, :n].cpu().numpy())


@pytest.mark.parametrize("mode", [S.CHOL_BF16X3, S.CHOL_F32])
def test_noalloc_factor_equals_allocating_entry(mode):
    for beads, extra in ((1100, 64), (330, 0)):  # order 3300, then the smaller 990
        sysm = S.synth_chain(beads, seed=beads)
        D = bench.rpy_matrix(sysm.coords, 3.8)
        A = D.astype(np.float32)
        n = A.shape[0]
        ref, _ = S.chol_factor(A, mode)
        got = _factor_in_place(n, A, mode, extra)
        assert np.isfinite(got).all()
        np.testing.assert_array_equal(got, ref, err_msg="order %d" % n)


def test_noalloc_factor_rejects_bad_sizes():
    import torch

    buf = torch.zeros((64, 64), dtype=torch.float32, device="cuda")
    f = _noalloc()
    assert f(0, MAX_ORDER + 1, buf.data_ptr(), MAX_ORDER + 64, S.CHOL_F32, None) == -1
    assert b"exceeds the maximum" in S.lib.sop_last_error()
    assert f(0, 100, buf.data_ptr(), 64, S.CHOL_F32, None) == -1  # ld below the padded order 128
    assert b"leading dimension" in S.lib.sop_last_error()
    assert f(0, 60, buf.data_ptr(), 66, S.CHOL_F32, None) == -1  # rows off 16-byte boundaries
    assert b"16-byte" in S.lib.sop_last_error()
    assert f(0, 64, buf.data_ptr(), 64, 7, None) == -1
    assert f(0, 0, buf.data_ptr(), 64, S.CHOL_F32, None) == -1
