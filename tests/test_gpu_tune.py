"""Autotuning on the GPU: fmm!(...; tune=true) and tune_fmm (src/fmm.jl:1073-1089,
src/autotune.jl:53-232).

The tuned values come from measured times, so the checks are the properties the reference's
documentation states (docs/src/advanced_usage_2.md:62-91): the returned expansion order is the
largest one the error-controlled M2L used, the returned leaf size follows from the evaluation's own
phase times, and fmm! with the tuned parameters meets the requested tolerance against the FP64
direct sum."""
import math

import numpy as np
import pytest

import fastmultipole_amd as F
from fastmultipole_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_fmm_tune_returns_order_used_and_balanced_leaf_size():
    print("[gpu] fmm tune=true", flush=True)
    sys_ = F.Gravitational(W.uniform_sources(20000, seed=31, strength_scale=1e-3))
    ev = F.Evaluation()
    optargs, cache, *_, ok = F.fmm(sys_, expansion_order=20, leaf_size=20, multipole_acceptance=0.5,
                                   error_tolerance=F.PowerAbsolutePotential(1e-6), scalar_potential=True,
                                   gradient=False, tune=True, update_target_systems=False, evaluation=ev)
    st = cache.last_stats
    assert ok and optargs["expansion_order"] == st.pmax == int(cache.m2l_orders().max())
    ratio = (st.ms_m2l / st.n_m2l) / (st.ms_p2p / st.n_p2p_pairs)
    assert optargs["leaf_size_source"] == math.ceil(math.sqrt(ratio))
    # without tune the arguments come back unchanged (src/fmm.jl:1160-1165)
    optargs2, *_ = F.fmm(sys_, expansion_order=7, leaf_size=33, multipole_acceptance=0.5,
                         update_target_systems=False, evaluation=ev)
    assert optargs2 == {"leaf_size_source": 33, "expansion_order": 7, "multipole_acceptance": 0.5}
    # farfield off: an empty M2L list halves the leaf size
    optargs3, *_ = F.fmm(sys_, expansion_order=4, leaf_size=33, farfield=False, tune=True,
                         update_target_systems=False, evaluation=ev)
    assert optargs3["leaf_size_source"] == 16


def test_tune_fmm_meets_tolerance():
    """The documentation's example (docs/src/advanced_usage_2.md:62-86) on 20,000 uniform sources
    (strengths U[0, 1e-3], potentials of order one)."""
    print("[gpu] tune_fmm", flush=True)
    bodies = W.uniform_sources(20000, seed=32, strength_scale=1e-3)
    system = F.Gravitational(bodies)
    tol = F.PowerAbsolutePotential(1e-4)
    params, cache = F.tune_fmm(system, error_tolerance=tol, scalar_potential=True, gradient=False, hessian=False,
                               verbose=False)
    assert params["multipole_acceptance"] in [round(0.3 + 0.1 * k, 10) for k in range(6)]
    assert 1 <= params["expansion_order"] <= 20 and params["leaf_size_source"] >= 1
    ref = F.Gravitational(bodies)
    F.direct(ref, scalar_potential=True, gradient=False)
    system.potential[:] = 0.0
    _, _, _, _, _, _, _, ok = F.fmm(system, scalar_potential=True, gradient=False, error_tolerance=tol,
                                    evaluation=cache, leaf_size=params["leaf_size_source"],
                                    expansion_order=params["expansion_order"],
                                    multipole_acceptance=params["multipole_acceptance"])
    assert ok
    err = np.abs(system.potential[:, 0] - ref.potential[:, 0]).max()
    assert err <= 1e-4, (params, err)


def test_autotune_measured_search_meets_tolerance():
    """core.autotune (the project's own tuner): every candidate is measured on the device, the chosen
    setting is the fastest measured one among those meeting the tolerance, its launch shape is applied
    to the context, and fmm! at the chosen setting meets the tolerance against the FP64 direct sum."""
    print("[gpu] autotune", flush=True)
    buf = W.vortex_ring(30000, seed=2)
    sys_ = F.VortexParticles(buf[:, 0:3], buf[:, 4:7], buf[:, 3])
    em = F.PowerRelativeGradient(1e-7)
    params, ev, hist = F.autotune(sys_, error_tolerance=em, expansion_order=20, leaf_sizes=(16, 32, 64),
                                  multipole_acceptances=(0.45, 0.55), free_slots=(0, 1), rounds=1, gradient=True)
    print(f"[gpu] autotune chose {params} from {len(hist)} candidates", flush=True)
    assert len(hist) >= 5 and all(h["ms"] > 0 for h in hist)
    feasible = [h for h in hist if h["error_success"]]
    assert feasible and params["error_success"]
    assert params["ms"] == min(h["ms"] for h in feasible)
    assert params["leaf_size"] in (16, 32, 64) and params["multipole_acceptance"] in (0.45, 0.55)
    _, _, *_, ok = F.fmm(sys_, expansion_order=20, leaf_size=params["leaf_size"],
                         multipole_acceptance=params["multipole_acceptance"], error_tolerance=em, gradient=True,
                         update_target_systems=False, evaluation=ev)
    assert ok
    with pytest.raises(Exception):
        ev.ctx.set_p2p_free_slots(4)
