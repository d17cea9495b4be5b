"""The reference's inputs files read into this library's parameters (CPU only). Fixtures: four
inputs files of the reference's example cases (exec/<case>/prob/inputs, copied as data under
tests/golden/inputs/)."""
import os

import numpy as np
import pytest

from compas_amd import inputs, make_params

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "inputs")


def _case(name, dim, nphase=0):
    return inputs.case_parameters(inputs.read(os.path.join(GOLDEN, name + ".inputs")), dim, nphase)


def test_parmparse_syntax():
    t = inputs.parse('a.b = 1 2 3   # comment\n# whole line\nname = "two words" x\na.b = 4\n\nbad line\n')
    assert t["a.b"] == ["4"]                       # the later definition replaces the earlier one
    assert t["name"] == ["two words", "x"]
    assert "bad line" not in t


def test_shock_bubble_air_helium():
    c = _case("ShockBubble-AirHelium-5Eq", 2)
    P, G, run = c["params"], c["geom"], c["run"]
    assert (P.id_weno, P.id_limiter, P.id_poly, P.id_bound) == (0, 1, 1, 0)        # "MUSCL"
    assert P.id_riemann == 2 and P.source_term == 1 and P.id_thinc == 0
    ref = make_params(2, gamma=(1.4, 1.6667), scheme="MUSCL", riemann="HLLC", source_term=1)
    for k in range(2):
        assert P.eos_A[k] == ref.eos_A[k] and P.eos_B[k] == ref.eos_B[k]
    assert list(G.domain_lo[:2]) == [0, 0] and list(G.domain_hi[:2]) == [95, 15]
    assert np.allclose(G.dx[:2], [3.0 / 96, 0.5 / 16])
    assert list(G.lo_bc[:2]) == [2, 4] and list(G.hi_bc[:2]) == [2, 4]
    assert run["cfl"] == 0.3 and run["timestep_change_limiter"] == 0 and run["max_level"] == 6
    assert c["pf"] is None
    # a six-level hierarchy is run without subcycling: only the subcycling setting is flagged
    assert not any("max_level" in u for u in c["unsupported"])
    assert any("do_subcycle" in u for u in c["unsupported"])
    assert run["ref_ratio"] == [2] * 6 and run["regrid_int"] == 1 and run["n_error_buf"] == 1
    assert run["blocking_factor"] == 8 and run["max_grid_size"] == 8
    assert not any("THINC" in u for u in c["unsupported"])     # THINC options count only with THINC = 1


def test_gas_impact_thinc_on_muscl_is_flagged():
    c = _case("GasImpact-5Eq", 2)
    assert c["params"].id_thinc == 1
    assert any("THINC = 1 on MUSCL" in u for u in c["unsupported"])
    assert list(c["geom"].lo_bc[:2]) == [2, 3]


def test_kelvin_helmholtz_weno5_periodic():
    c = _case("KelvinHelmholtz-5Eq", 2)
    P, G = c["params"], c["geom"]
    assert (P.id_weno, P.id_limiter, P.id_poly) == (1, 5, 4) and c["run"]["ghost_cells"] == 3
    assert list(G.is_periodic[:2]) == [1, 0]
    assert not any("THINC" in u for u in c["unsupported"])


def test_three_material_nphase():
    c = _case("ThreeMaterialRM-5Eq-N", 2, nphase=3)
    P = c["params"]
    assert P.nphase == 3 and P.nstate == 9
    assert (P.id_weno, P.id_limiter, P.id_poly) == (0, 4, 1)                          # "MUSCL-MC"
    for k, g in enumerate((1.1074, 1.2040, 1.400)):
        assert P.eos_A[k] == 1.0 / (g - 1.0)
    assert P.source_term == 1


def test_phase_field_keys():
    text = """
    amr.n_cell = 32 32
    Physics.source_term = 1
    FiniteVolume.Scheme = "WENO5"
    prob.EOS.EOS = 1
    prob.EOS.gamma_1 = 4.4
    prob.EOS.gamma_2 = 1.4
    prob.EOS.pinf_1 = 6.0e8
    prob.EOS.pinf_2 = 0.0
    PhaseField.phase_field = 1
    PhaseField.Mechanism = "ACDI"
    PhaseField.ID_ExplicitRC = 1
    PhaseField.cfl = 0.3
    PhaseField.Eta_Multiplier = 0.25
    """
    c = inputs.case_parameters(inputs.parse(text), 2)
    pf, P = c["pf"], c["params"]
    assert pf.mechanism == 3 and pf.explicit_rc == 1 and pf.cfl == 0.3 and pf.eta_multiplier == 0.25
    assert pf.flux_phase2flux == 1 and pf.eta_factor == 2.0
    # p_min = -B / (A + 1) with the source term (source/Parm.cpp:255-262)
    assert pf.p_min[0] == -P.eos_B[0] / (P.eos_A[0] + 1.0) and pf.p_min[1] == 0.0
    # CAC-Adv with the explicit RC is refused, as the reference aborts
    with pytest.raises(ValueError):
        inputs.case_parameters(inputs.parse(text.replace('"ACDI"', '"CAC-Adv"')), 2)
    # the default mechanism is CAC-Adv with the implicit RC
    c = inputs.case_parameters(inputs.parse("amr.n_cell = 8 8\nPhaseField.phase_field = 1\n"), 2)
    assert c["pf"].mechanism == 1 and c["pf"].explicit_rc == 0


def test_time_integrator_key():
    from compas_amd import inputs
    base = "amr.n_cell = 16 16\n"
    assert inputs.case_parameters(inputs.parse(base), 2)["run"]["time_integrator"] == 3
    for name, code in (("ForwardEuler", 1), ("TVD-RK2", 2), ("TVD-RK3", 3), ("RK4", 4)):
        t = inputs.parse(base + 'run.TimeIntegrator = "%s"\n' % name)
        assert inputs.case_parameters(t, 2)["run"]["time_integrator"] == code
    with pytest.raises(ValueError):
        inputs.case_parameters(inputs.parse(base + "run.TimeIntegrator = RK5\n"), 2)
