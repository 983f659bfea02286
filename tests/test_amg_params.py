"""AMG parameters (host code, no GPU): the defaults are the reference's, and
values the V-cycle cannot use are refused before the library touches a
device (the refusal returns UNAP_HIP_ERR_ARG, not a device error)."""
import pytest

from unap_amd import hip, meshgen

ERR_ARG = -1


def test_defaults_are_the_reference():
    p = hip.amg_params()
    # MG.cpp:11-13, MGPreconditioner.cpp:11, lduAgglomeration.cpp:11 and the
    # Chebyshev interval of chebySmoother.hpp:46-47
    assert (p.smoother, p.n_cells_coarsest, p.max_levels, p.n_vcycles) == (1, 50, 50, 2)
    assert (p.n_pre_sweeps, p.n_post_sweeps, p.n_finest_sweeps) == (0, 2, 2)
    assert (p.precision, p.level_sweep_growth, p.pair_steps, p.cycle, p.scaling) == (64, 1, 1, hip.CYCLE_AUTO, 2)
    assert (p.cheby_eig_ratio, p.cheby_boost) == (1.5, 1.1)
    assert p.cheby_eig_ratio_finest == 0.0  # the one ratio on every level


@pytest.mark.parametrize("kw", [dict(cheby_eig_ratio=1.0), dict(cheby_eig_ratio=0.5), dict(cheby_boost=0.0),
                                dict(cheby_boost=float("nan")), dict(cheby_eig_ratio=float("inf")),
                                dict(cheby_eig_ratio_finest=1.0), dict(cheby_eig_ratio_finest=-2.0),
                                dict(cheby_eig_ratio_finest=float("nan")),
                                dict(scaling=3), dict(pair_steps=0), dict(smoother=3)])
def test_bad_parameters_refused(kw):
    s = meshgen.config_system(4, 6)
    A = hip.LduMatrix.from_system(s)
    with pytest.raises(hip.UnapHipError) as e:
        hip.Precond(A, "AMG", amg=hip.amg_params(**{"smoother": 2, "precision": 32, "cycle": hip.CYCLE_FIXED, **kw}))
    assert e.value.code == ERR_ARG, e.value
