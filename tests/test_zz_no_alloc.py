"""The allocation-free MST entry point against the ordinary device entry
point, through the C-ABI: identical outputs on consecutive calls (the
second smaller), on a caller's stream, and a clear error for inputs beyond
its scratch."""
import ctypes

import numpy as np
import pytest

from fast_hdbscan_amd import workloads

pytestmark = pytest.mark.gpu


def _outputs(n, ms, dev):
    import torch

    kn = ms if ms > 1 else 1
    return {
        "core": torch.full((n,), -1.0, dtype=torch.float32, device=dev),
        "nbr": torch.full((n, kn), -7, dtype=torch.int32, device=dev),
        "src": torch.full((max(n - 1, 1),), -7, dtype=torch.int32, device=dev),
        "dst": torch.full((max(n - 1, 1),), -7, dtype=torch.int32, device=dev),
        "w": torch.full((max(n - 1, 1),), -1.0, dtype=torch.float32, device=dev),
    }


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _noalloc(X, ms):
    import torch

    from fast_hdbscan_amd import _lib

    n, d = X.shape
    o = _outputs(n, ms, X.device)
    st = _lib.FHStats()
    err = ctypes.create_string_buffer(512)
    stream = torch.cuda.current_stream(X.device)
    torch.cuda.synchronize()
    code = _lib.load().fh_mst_euclidean_device_noalloc(
        0, ctypes.c_void_p(stream.cuda_stream), _p(X), n, d, ms, _p(o["core"]), _p(o["nbr"]), _p(o["src"]),
        _p(o["dst"]), _p(o["w"]), ctypes.byref(st), err, 512)
    return code, err.value.decode(), o, st


def _ordinary(X, ms):
    from fast_hdbscan_amd import _lib

    n, d = X.shape
    o = _outputs(n, ms, X.device)
    ctx = _lib.context(0)
    code = _lib.load().fh_mst_euclidean_device(ctx, _p(X), n, d, ms, None, _p(o["core"]), _p(o["nbr"]),
                                               _p(o["src"]), _p(o["dst"]), _p(o["w"]), None)
    _lib.check(ctx, code)
    return o


@pytest.mark.parametrize("order", ["large_then_small"])
def test_no_alloc_matches_device_entry(order):
    import torch

    cases = [("2-D scale workload", workloads.scale_2d(300_000, seed=31), 10),
             ("blobs", workloads.blobs_small(40_000, seed=32), 5),
             ("20-D blobs (MFMA path)", np.random.default_rng(33).standard_normal((4000, 20)).astype(np.float32), 4)]
    for name, Xh, ms in cases:
        print(f"no-alloc case {name} n={Xh.shape[0]} d={Xh.shape[1]} ms={ms}", flush=True)
        X = torch.from_numpy(Xh).cuda()
        code, msg, got, st = _noalloc(X, ms)
        assert code == 0, msg
        want = _ordinary(X, ms)
        torch.cuda.synchronize()
        for k in ("core", "nbr", "src", "dst", "w"):
            np.testing.assert_array_equal(got[k].cpu().numpy(), want[k].cpu().numpy(), err_msg=k)
        assert st.n_edges == Xh.shape[0] - 1


def test_no_alloc_rejects_inputs_beyond_its_scratch():
    import torch

    X = torch.from_numpy(workloads.scale_2d(2_000_000, seed=34)).cuda()
    print("no-alloc case n=2000000 (beyond the scratch)", flush=True)
    code, msg, got, _ = _noalloc(X, 10)
    assert code == -1 and "scratch" in msg
    # nothing was launched: the outputs keep their fill
    assert int((got["src"] != -7).sum()) == 0
    # and the entry point still works afterwards
    Xs = torch.from_numpy(workloads.blobs_small(5000, seed=35)).cuda()
    code, msg, got, _ = _noalloc(Xs, 10)
    assert code == 0, msg
    want = _ordinary(Xs, 10)
    np.testing.assert_array_equal(got["src"].cpu().numpy(), want["src"].cpu().numpy())
