"""The allocation-free FoF entry point (nbx_fof_noalloc, static device
scratch) against nbx_fof through a tree: bit-identical labels on calls in
a row (the second smaller, so the scratch holds the first call's data),
at the maximum size, on a caller's stream, and the size limit."""
import numpy as np
import pytest
import torch

import pyoracle as O
from helpers import clustered, disk

pytestmark = pytest.mark.gpu


def _labels_tree(nbx, p, ll):
    t = nbx.allocateOctree(p.n)
    out = nbx.fof_labels(t, p, ll).cpu().numpy()
    t.free()
    return out


def test_noalloc_fof_equals_tree_fof_twice(gpu):
    nbx = gpu
    cases = [(disk(60000, seed=81), 0.0261), (clustered(900, 5, 1e-3, seed=82), 2e-3),
             (disk(3000, seed=83), 0.05)]
    for d, ll in cases:
        n = len(d["x"])
        print(f"noalloc fof n={n} ll={ll}", flush=True)
        p = nbx.to_gpu(d)
        lab = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        nbx.fof_noalloc(p, ll, lab)
        got = lab.cpu().numpy()
        np.testing.assert_array_equal(got, _labels_tree(nbx, p, ll))
        ref, _ = O.fof(d["x"], d["y"], d["z"], ll)
        np.testing.assert_array_equal(got, ref)


def test_noalloc_fof_max_size_and_stream(gpu):
    nbx = gpu
    n = nbx.NOALLOC_MAX_BODIES
    d = disk(n, seed=84)
    p = nbx.to_gpu(d)
    lab = torch.empty(n, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    print(f"noalloc fof n={n} on a caller stream", flush=True)
    nbx.fof_noalloc(p, 0.0261, lab, stream=s)
    np.testing.assert_array_equal(lab.cpu().numpy(), _labels_tree(nbx, p, 0.0261))
    d2 = disk(n + 1, seed=85)
    p2 = nbx.to_gpu(d2)
    lab2 = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="exceed"):
        nbx.fof_noalloc(p2, 0.0261, lab2)
    # a smaller call after the refused one still matches
    d3 = disk(777, seed=86)
    p3 = nbx.to_gpu(d3)
    lab3 = torch.empty(777, dtype=torch.int32, device="cuda")
    nbx.fof_noalloc(p3, 0.05, lab3)
    np.testing.assert_array_equal(lab3.cpu().numpy(), _labels_tree(nbx, p3, 0.05))
