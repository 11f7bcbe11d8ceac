"""Block-step schedule identities of the mutual lane (CPU; the reference's
force.py:685-882 and tests/integration/test_mutual_fmm.py:428-454)."""

from __future__ import annotations

import numpy as np
import pytest

from jaccpot_amd import mutual as mu


def test_sub_steps_and_sync_boundaries():
    assert [mu.n_sub(k) for k in range(4)] == [1, 2, 4, 8]
    assert [s for s in range(9) if mu.is_sync_boundary(s, 3)] == [0, 8]


def test_active_floor_follows_trailing_zeros():
    # boundary s is reached by the levels whose step 2**(k_max-k) divides s
    assert [mu.active_level_floor(s, 3) for s in range(9)] == [0, 3, 2, 3, 1, 3, 2, 3, 0]
    for k_max in range(1, 6):
        for s in range(1, mu.n_sub(k_max)):
            f = mu.active_level_floor(s, k_max)
            assert s % (1 << (k_max - f)) == 0 and s % (1 << (k_max - f + 1)) != 0


def test_weights_and_table():
    w = mu.level_weights_from_floor(1, 3, 0.08, half=0.5)
    assert w.tolist() == [0.0, 0.02, 0.01, 0.005]
    tab = mu.boundary_weight_table(2, 0.1)
    assert tab.shape == (5, 3)
    assert np.allclose(tab[0], [0.05, 0.025, 0.0125]) and np.allclose(tab[4], tab[0])
    assert np.allclose(tab[1], [0, 0, 0.025]) and np.allclose(tab[2], [0, 0.05, 0.025])
    # over a base step every level is kicked by dt_max / 2**k per sub-step of its own, in total dt_max
    k_max, dt = 3, 0.2
    t = mu.boundary_weight_table(k_max, dt)
    assert np.allclose(t.sum(0), dt)


def test_block_step_model_rejects_bad_arguments():
    with pytest.raises(ValueError):
        mu.BlockStepFMM(softening=0.01, k_max=-1)
    with pytest.raises(ValueError):
        mu.BlockStepFMM(softening=0.01, k_max=2, basis="complex")
    with pytest.raises(ValueError):
        mu.BlockStepFMM(softening=0.01, k_max=2, backend="triton")
