"""The reference's keyword surface (jaccpot/solver.py:664-692, 1012-1024,
1358-1368; jaccpot/config.py:420-554) resolved on the host: execution knobs are
accepted, algorithm changes this path does not implement raise
NotImplementedError, out-of-domain values ValueError, unknown names TypeError."""

from __future__ import annotations

import pytest

from jaccpot_amd.config import (FarFieldConfig, FMMAdvancedConfig, GradConfig, TreeConfig,
                                resolve_constructor_options, resolve_grad_config, resolve_prepare_options)

DEFAULTS = dict(adaptive_order=False, use_pallas=None, mac_force_scale_mode="prev",
                mac_force_scale_prepass_theta=None, mac_force_scale_fb_inflation=1.0,
                adaptive_error_model="tail_proxy", adaptive_eps=None, dehnen_geometry_mode="com",
                mac_theta_max=1.0, advanced=None, legacy_kwargs={})


def resolve(**kw):
    return resolve_constructor_options(**{**DEFAULTS, **kw})


def test_defaults_resolve():
    opts = resolve()
    assert opts["mac_force_scale_mode"] == "prev" and opts["inert_legacy"] == []


@pytest.mark.parametrize("kw", [
    dict(use_pallas=True), dict(use_pallas=False),
    # adaptive-acceptance knobs are inert without adaptive_order (as in the reference)
    dict(mac_force_scale_mode="paper_fb"), dict(adaptive_error_model="dehnen_paper"), dict(adaptive_eps=1e-4),
    dict(dehnen_geometry_mode="exact"), dict(mac_theta_max=0.8), dict(mac_force_scale_prepass_theta=0.5),
    dict(advanced=FMMAdvancedConfig()),
    dict(advanced=FMMAdvancedConfig(farfield=FarFieldConfig(m2l_chunk_size=4096, retain_far_pairs_for_grad=True))),
    dict(legacy_kwargs={"m2l_chunk_size": 8192, "retain_far_pairs_for_grad": True, "fail_fast": True,
                        "memory_objective": "min", "nearfield_mode": "bucketed", "runtime_path": "auto"}),
    dict(legacy_kwargs={"expansion_basis": "real"}),
    dict(legacy_kwargs={"mac_type": "dehnen", "tree_type": "radix", "refine_local": False}),
])
def test_accepted(kw):
    resolve(**kw)


@pytest.mark.parametrize("kw", [
    dict(mac_force_scale_mode="bogus"), dict(adaptive_error_model="bogus"), dict(dehnen_geometry_mode="bogus"),
    dict(mac_theta_max=0.0), dict(mac_theta_max=1.5), dict(adaptive_eps=-1.0), dict(mac_force_scale_prepass_theta=2.0),
    dict(mac_force_scale_fb_inflation=-0.1), dict(use_pallas="yes"),
])
def test_out_of_domain_raises_value_error(kw):
    with pytest.raises(ValueError):
        resolve(**kw)


@pytest.mark.parametrize("kw", [
    dict(adaptive_order=True, adaptive_eps=1e-4),
    dict(adaptive_order=True, dehnen_geometry_mode="exact"),
    dict(adaptive_order=True, adaptive_error_model="dehnen_paper"),
    dict(adaptive_order=True, mac_force_scale_mode="paper"),
    dict(adaptive_order=True, mac_theta_max=0.9),
    dict(advanced=FMMAdvancedConfig(tree=TreeConfig(refine_local=True))),
    dict(advanced=FMMAdvancedConfig(farfield=FarFieldConfig(grouped_interactions=True))),
    dict(advanced=FMMAdvancedConfig(farfield=FarFieldConfig(mixed_order=True))),
    dict(advanced=FMMAdvancedConfig(mac_type="bh")),
    dict(advanced=FMMAdvancedConfig(dehnen_radius_scale=1.2)),
    dict(legacy_kwargs={"tree_type": "octree"}),
    dict(legacy_kwargs={"grouped_interactions": True}),
    dict(legacy_kwargs={"fixed_order": 4}),
    dict(legacy_kwargs={"target_leaf_particles": 32}),
])
def test_algorithm_changes_raise_not_implemented(kw):
    with pytest.raises(NotImplementedError):
        resolve(**kw)


def test_adaptive_defaults_accepted():
    resolve(adaptive_order=True)


def test_unknown_keyword_raises_type_error():
    with pytest.raises(TypeError, match="no_such_option"):
        resolve(legacy_kwargs={"no_such_option": 1})


def test_prepare_options():
    resolve_prepare_options(adaptive_order=False, cache_policy="auto", force_scale_nodes=None,
                            runtime_overrides_override=None, fused_device_mode=True,
                            extra={"retain_far_pairs_for_grad": True})
    resolve_prepare_options(adaptive_order=False, cache_policy="reuse", force_scale_nodes=[1.0],
                            runtime_overrides_override=object(), fused_device_mode=False, extra={})
    with pytest.raises(NotImplementedError):
        resolve_prepare_options(adaptive_order=True, cache_policy="auto", force_scale_nodes=[1.0],
                                runtime_overrides_override=None, fused_device_mode=False, extra={})
    with pytest.raises(TypeError):
        resolve_prepare_options(adaptive_order=False, cache_policy="auto", force_scale_nodes=None,
                                runtime_overrides_override=None, fused_device_mode=False, extra={"bogus": 1})


def test_grad_config():
    cfg = GradConfig(nearfield_lane="fast_lane", fused_m2l_pallas=True)  # README.md:533-534
    assert resolve_grad_config(cfg, None) is cfg
    assert resolve_grad_config(None, None) is None
    with pytest.raises(ValueError):
        GradConfig(nearfield_lane="bogus")
    with pytest.raises(ValueError):
        GradConfig(nearfield_fast_lane_min_particles=-1)
    with pytest.raises(NotImplementedError):
        resolve_grad_config(None, object())

    class Duck:  # the reference's own GradConfig instance has the same fields
        nearfield_lane = "bucketed"
        reverse_tiers = 2

    got = resolve_grad_config(Duck(), None)
    assert got.nearfield_lane == "bucketed" and got.reverse_tiers == 2
