"""Loading a prepared state checks every index array before it reaches a
kernel: a corrupted or truncated file raises ValueError on the host."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from jaccpot_amd.persist import validate_topology
from jaccpot_amd.topology import build_topology
from oracle import ics


@pytest.fixture(scope="module")
def topo():
    x, m = ics.plummer(3000, 3)
    return build_topology(x, m, leaf_size=16, theta=0.6)


@pytest.mark.parametrize("n,leaf", [(1, 16), (7, 4), (777, 5), (3, 256), (4096, 16)])
def test_valid_topologies_pass(n, leaf):
    x, m = ics.uniform(n, 0)
    validate_topology(build_topology(x, m, leaf_size=leaf, theta=0.6))


def _bad(topo, **changes):
    return dataclasses.replace(topo, **changes)


def _corruptions(t):
    perm = t.perm.copy()
    perm[0] = perm[1]
    far_off = t.far_off.copy()
    far_off[5] = far_off[6] + 1  # not monotone
    far_src = t.far_src.copy()
    far_src[0] = t.num_nodes
    near_src = t.near_src.copy()
    near_src[-1] = -1
    rng = t.node_range.copy()
    rng[3, 1] = t.num_particles
    parent = t.parent.copy()
    parent[0] = 0
    left = t.left.copy()
    left[2] = 0
    return {
        "duplicate in perm": dict(perm=perm),
        "truncated perm": dict(perm=t.perm[:-1]),
        "far offsets not monotone": dict(far_off=far_off),
        "far source out of range": dict(far_src=far_src),
        "far list truncated": dict(far_src=t.far_src[:-3]),
        "negative near source": dict(near_src=near_src),
        "node range past N": dict(node_range=rng),
        "root with a parent": dict(parent=parent),
        "child pointing at the root": dict(left=left),
        "int64 indices": dict(far_src=t.far_src.astype(np.int64)),
        "node counts": dict(num_leaves=t.num_leaves + 1),
        "m2l target out of range": dict(m2l_targets=np.append(t.m2l_targets, np.int32(t.num_nodes))),
        "non-finite centre": dict(center=np.where(np.arange(t.num_nodes)[:, None] == 4, np.nan, t.center)),
    }


CORRUPTIONS = ["duplicate in perm", "truncated perm", "far offsets not monotone", "far source out of range",
               "far list truncated", "negative near source", "node range past N", "root with a parent",
               "child pointing at the root", "int64 indices", "node counts", "m2l target out of range",
               "non-finite centre"]


@pytest.mark.parametrize("name", CORRUPTIONS)
def test_corrupted_topology_raises(topo, name):
    with pytest.raises(ValueError):
        validate_topology(_bad(topo, **_corruptions(topo)[name]))
