"""GPU: the device topology build against the host builder, array by array,
bit for bit (tree, geometry, level schedules, far and near lists)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import ics

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FIELDS = ["perm", "parent", "left", "right", "node_range", "depth", "center", "radius", "mass", "far_off", "far_src",
          "near_off", "near_src", "m2m_nodes", "m2m_off", "l2l_nodes", "l2l_off", "m2l_targets"]


def _inputs(kind, n, seed):
    if kind == "duplicates":  # many exactly coincident particles: equal Morton codes, index tie-break
        x, m = ics.uniform(n, seed)
        return np.round(x * 4) / 4, m
    if kind == "massless":  # every node takes the plain-mean centre
        x, m = ics.uniform(n, seed)
        return x, np.zeros_like(m)
    if kind == "some_massless":
        x, m = ics.plummer(n, seed)
        m = m.copy()
        m[: n // 2] = 0.0
        return x, m
    return getattr(ics, kind)(n, seed)


CASES = [
    # (kind, n, leaf, theta, bounds)
    ("uniform", 4096, 16, 0.6, None),
    ("plummer", 20000, 8, 0.85, None),
    ("blobs", 10000, 64, 0.6, None),
    ("uniform", 7, 4, 0.6, None),
    ("uniform", 1, 16, 0.6, None),
    ("uniform", 1000, 1, 0.5, None),
    ("plummer", 5000, 256, 0.7, None),
    ("duplicates", 3000, 8, 0.6, None),
    ("massless", 2000, 8, 0.6, None),
    ("some_massless", 3000, 16, 0.7, None),
    ("uniform", 3000, 16, 0.6, ((-2.0, -2.0, -2.0), (3.0, 3.0, 3.0))),
    ("plummer", 262144, 8, 0.85, None),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}{c[1]}_leaf{c[2]}_th{c[3]}{'_bounds' if c[4] else ''}"
                                             for c in CASES])
def test_device_build_matches_host(case, require_gpu):
    from jaccpot_amd.dtopology import build_topology_device
    from jaccpot_amd.topology import build_topology

    kind, n, leaf, theta, bounds = case
    print(f"[gpu] device topology {kind} n={n} leaf={leaf} theta={theta}", flush=True)
    x, m = _inputs(kind, n, 3)
    host = build_topology(x, m, leaf_size=leaf, theta=theta, bounds=bounds)
    dev = build_topology_device(torch.as_tensor(x, device="cuda"), torch.as_tensor(m, device="cuda"),
                                leaf_size=leaf, theta=theta, bounds=bounds)
    for k in ("num_particles", "num_leaves", "num_internal", "num_nodes", "max_depth"):
        assert getattr(dev, k) == getattr(host, k), k
    for f in FIELDS:
        a, b = getattr(dev, f), getattr(host, f)
        assert a.shape == b.shape, (f, a.shape, b.shape)
        assert np.array_equal(a, b), f
    print(f"[gpu]   {host.num_far_pairs} far pairs, {host.num_near_entries} near entries", flush=True)


def test_device_build_rejects_non_finite(require_gpu):
    from jaccpot_amd.dtopology import build_topology_device

    x, m = ics.uniform(100, 1)
    x[17, 1] = np.nan
    with pytest.raises(ValueError):
        build_topology_device(torch.as_tensor(x, device="cuda"), torch.as_tensor(m, device="cuda"), leaf_size=8,
                              theta=0.6)
