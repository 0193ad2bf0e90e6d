"""pamo_point_mesh_sqdist_noalloc (static scratch, no allocation) against
pamo_point_mesh_sqdist (context buffers), through the C ABI: two calls in a
row, the second on a smaller mesh and fewer points, with squared distances,
face indices and closest points equal bit for bit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _case(level, seed, n_points):
    from pamo_amd import synth
    V, F = synth.icosphere_sh(level, seed)
    rng = np.random.default_rng(seed)
    P = rng.uniform(-1.3, 1.3, (n_points, 3))
    return P, V.astype(np.float64), F.astype(np.int32)


def _run(torch, fn, P, V, F, ctx_ptr=None):
    from pamo_amd import _hip
    Pd = torch.from_numpy(P).cuda()
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(F).cuda()
    n = len(P)
    sq = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
    idx = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    cp = torch.full((n, 3), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    lib = _hip.lib()
    if fn == "ctx":
        rc = lib.pamo_point_mesh_sqdist(ctx_ptr, Pd.data_ptr(), n, Vd.data_ptr(), len(V), Fd.data_ptr(), len(F),
                                        sq.data_ptr(), idx.data_ptr(), cp.data_ptr())
    else:
        stream = torch.cuda.current_stream().cuda_stream
        rc = lib.pamo_point_mesh_sqdist_noalloc(Pd.data_ptr(), n, Vd.data_ptr(), len(V), Fd.data_ptr(), len(F),
                                                sq.data_ptr(), idx.data_ptr(), cp.data_ptr(), stream)
    _hip.check(rc, fn)
    torch.cuda.synchronize()
    return sq.cpu().numpy(), idx.cpu().numpy(), cp.cpu().numpy()


def test_noalloc_matches_context_entry_point_twice(torch_cuda):
    from pamo_amd import _hip
    torch = torch_cuda
    m = ctypes.c_void_p()
    _hip.check(_hip.lib().pamo_metrics_create(ctypes.byref(m), 0), "pamo_metrics_create")
    try:
        # 20,480 faces then 1,280 faces: the second call reuses the scratch
        # the first one filled
        for level, seed, n_points in ((5, 1, 50000), (3, 2, 7000)):
            P, V, F = _case(level, seed, n_points)
            print(f"case: no-alloc point-mesh, {len(F)} faces, {n_points} points", flush=True)
            a = _run(torch, "ctx", P, V, F, m)
            b = _run(torch, "noalloc", P, V, F)
            assert np.array_equal(a[0].view(np.uint64), b[0].view(np.uint64))
            assert np.array_equal(a[1], b[1])
            assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
            assert np.all(a[1] >= 0) and np.all(a[0] >= 0)
    finally:
        _hip.lib().pamo_metrics_destroy(m)


def test_noalloc_rejects_meshes_over_its_limit(torch_cuda):
    """More faces than the static scratch holds: an error status before any
    device work."""
    from pamo_amd import _hip
    torch = torch_cuda
    P, V, F = _case(7, 3, 10)  # 327,680 faces
    with pytest.raises(_hip.PamoError, match="static scratch"):
        _run(torch, "noalloc", P, V, F)
