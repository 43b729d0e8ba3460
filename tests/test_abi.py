"""C-ABI surface and host-side logic of libfh.so (CPU only: no device calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from fast_hdbscan_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "fh.h")).read()
    return sorted(set(re.findall(r"\b(fh_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.load()
    names = _declared()
    assert "fh_mst_euclidean_device" in names
    for name in names:
        assert hasattr(L, name), f"{name} declared in include/fh.h but not exported"
    assert set(_lib.EXPORTED_SYMBOLS) == set(names)


def test_version_string():
    assert _lib.load().fh_version().decode().startswith("fh ")


def test_tree_nnodes_closed_form():
    L = _lib.load()
    for n in [1, 2, 41, 81, 82, 10_000, 2_000_000]:
        levels = int(np.log2(max(1, (n - 1) / 40)) + 1)
        assert L.fh_tree_nnodes(n) == 2 ** levels - 1


def test_null_context_is_rejected():
    L = _lib.load()
    X = np.zeros((4, 2), np.float32)
    out = np.zeros(4, np.float32)
    code = L.fh_mst_euclidean(None, _lib.ptr(X), 4, 2, 2, None, _lib.ptr(out), None, None, None, None, None)
    assert code == _lib.FH_EINVAL


@pytest.mark.parametrize("n,d", [(1, 2), (500, 2), (20_000, 3), (5_000, 5)])
def test_host_tree_build_matches_oracle(n, d):
    from oracle import oracle as O
    from fast_hdbscan_amd.hdbscan import build_kdtree

    X = np.random.default_rng(n).standard_normal((n, d)).astype(np.float32)
    t = build_kdtree(X)  # host builder (no device)
    o = O.build_kdtree(X)
    for k in ("idx_array", "idx_start", "idx_end", "node_bounds"):
        np.testing.assert_array_equal(t[k], getattr(o, k))
    np.testing.assert_array_equal(t["is_leaf"], o.is_leaf)


def test_label_stage_rejects_bad_params():
    from fast_hdbscan_amd.labels import label_params, labels_from_sorted_arrays

    with pytest.raises(ValueError):
        label_params(10, cluster_selection_method="fail")
    p = label_params(0)
    with pytest.raises(ValueError):
        labels_from_sorted_arrays(3, np.array([0, 1]), np.array([1, 2]), np.array([1.0, 2.0], np.float32), p)


@pytest.mark.parametrize("struct", ["FHStats", "FHLabelParams", "FHLabelOrder", "FHCondensedTree", "FHLayerParams",
                                    "FHLayers"])
def test_ctypes_structs_match_the_header(tmp_path, struct):
    """Every ctypes mirror has the C struct's size and field offsets
    (gcc compiles a probe against include/fh.h)."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    cls = getattr(_lib, struct)
    cname = {"FHStats": "fh_stats", "FHLabelParams": "fh_label_params", "FHLabelOrder": "fh_label_order",
             "FHCondensedTree": "fh_condensed_tree", "FHLayerParams": "fh_layer_params",
             "FHLayers": "fh_layers"}[struct]
    fields = [f for f, _ in cls._fields_]
    src = tmp_path / "probe.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fh.h"', "int main(void) {",
             f'  printf("%zu\\n", sizeof({cname}));']
    lines += [f'  printf("%zu\\n", offsetof({cname}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == ctypes.sizeof(cls)
    for f, off in zip(fields, out[1:]):
        assert getattr(cls, f).offset == off, f
