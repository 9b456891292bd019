"""The CTI reader (hpo_mech_load on a .cti file): the reference ships its mechanisms as CTI
beside YAML and CTML (resources/mechanisms/{gri30,h2o2,nDodecane_Reitz}.cti, copied unchanged
into hpc_ode_amd/mechanisms/). Cantera's YAML files were converted from the same sources, so
the tables read from the CTI file must equal those read from the YAML file (which the oracle's
reader and the pyJac golden vectors pin): species order, molecular weights, NASA7
coefficients, reaction kinds, reversibility and every Arrhenius parameter after the unit
conversion. tests/test_gpu_cti.py checks the rest of the tables (stoichiometry,
efficiencies, Troe parameters) through the RHS and Jacobian on the GPU."""
import os
import re

import numpy as np
import pytest

import hpc_ode_amd as H

MECH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hpc_ode_amd", "mechanisms")


@pytest.mark.parametrize("name", ["gri30", "nDodecane_Reitz"])
def test_cti_tables_equal_yaml_tables(name):
    c = H.Mechanism(os.path.join(MECH, name + ".cti"))
    y = H.Mechanism(os.path.join(MECH, name + ".yaml"))
    assert c.species == y.species
    assert (c.n_species, c.n_state, c.n_reactions) == (y.n_species, y.n_state, y.n_reactions)
    _same_tables(c, y, c.n_species)


def test_cti_h2o2_species():
    """The reference's h2o2.cti is an older edition of its h2o2.yaml: the nine H/O/Ar species
    (the YAML appends N2 as a tenth, h2o2.yaml:16) and 28 reactions (the YAML splits one more
    duplicate pair into 29). The species' data agree."""
    c = H.Mechanism(os.path.join(MECH, "h2o2.cti"))
    y = H.Mechanism(os.path.join(MECH, "h2o2.yaml"))
    assert c.species + ["N2"] == y.species
    assert (c.n_reactions, y.n_reactions) == (28, 29)
    _same_species(c, y, c.n_species)


def _same_species(c, y, K):
    assert np.array_equal(c.W[:K], y.W[:K])
    tc, ty = c.tables(), y.tables()
    assert np.array_equal(tc["tmid"][:K], ty["tmid"][:K])
    assert np.array_equal(tc["nasa"][:K], ty["nasa"][:K])


def _same_tables(c, y, K):
    _same_species(c, y, K)
    tc, ty = c.tables(), y.tables()
    assert np.array_equal(tc["kind"], ty["kind"])
    assert np.array_equal(tc["rev"], ty["rev"])
    # the same numbers converted by the same factors: equal to the last bit or to one rounding
    # of the unit conversion (A 1e-15, theta = E * 4184 / R)
    assert np.allclose(tc["arr"], ty["arr"], rtol=2e-15, atol=0)


def test_cti_unit_conversion_by_hand(tmp_path):
    """A two-species file in cm/mol/kcal: A of a bimolecular reaction scales by 1e-3
    (cm^3/mol -> m^3/kmol), of a three-body reaction by 1e-6, theta = E / R."""
    f = tmp_path / "tiny.cti"
    f.write_text('''units(length='cm', time='s', quantity='mol', act_energy='kcal/mol')
ideal_gas(name='tiny', elements='H', species='H H2', reactions='all',
          initial_state=state(temperature=300.0, pressure=OneAtm))
species(name='H', atoms='H:1',
        thermo=(NASA([200.0, 1000.0], [2.5, 0.0, 0.0, 0.0, 0.0, 25473.0, -0.44]),
                NASA([1000.0, 3500.0], [2.5, 0.0, 0.0, 0.0, 0.0, 25473.0, -0.44])))
species(name='H2', atoms='H:2',
        thermo=(NASA([200.0, 1000.0], [2.34, 0.008, -1.9e-05, 2.0e-08, -7.4e-12, -917.9, 0.68]),
                NASA([1000.0, 3500.0], [3.34, -4.9e-05, 5.0e-07, -1.8e-10, 2.0e-14, -950.2, -3.2])))
# comment line
reaction("H + H2 => H2 + H", [2.0e+13, 0.5, 1.5], options='duplicate')
three_body_reaction('2 H + M <=> H2 + M', [1.0e+18, -1.0, 0.0],
                    efficiencies = "H2:2.5")
''')
    m = H.Mechanism(str(f))
    assert m.species == ["H", "H2"]
    assert np.allclose(m.W, [1.008, 2.016], rtol=1e-15)
    t = m.tables()
    R = 8314.4621
    assert t["arr"][0, 3] == pytest.approx(2.0e13 * 1e-3, rel=1e-15)
    assert t["arr"][0, 1] == 0.5
    assert t["arr"][0, 2] == pytest.approx(1.5 * 4184.0e3 / R, rel=1e-15)
    assert t["rev"][0] == 0
    assert t["arr"][1, 3] == pytest.approx(1.0e18 * 1e-6, rel=1e-15)
    assert t["rev"][1] == 1
    assert list(t["kind"]) == [0, 1]


@pytest.mark.parametrize("text,err", [
    ("ideal_gas(name='x', species='A', reactions='all')\nspecies(name='A', atoms='H:1', thermo=NASA9([200, 1000], [1,2,3,4,5,6,7,8,9]))\n",
     "NASA7"),
    ("units(length='cm', quantity='mol', act_energy='cal/mol')\nideal_gas(name='x', species='H', reactions='all')\n"
     "species(name='H', atoms='H:1', thermo=(NASA([200.0, 1000.0], [2.5,0,0,0,0,25473.0,-0.44]),))\n"
     "chebyshev_reaction('H + H <=> H', Tmin=300.0, Tmax=2000.0, Pmin=(0.1, 'atm'), Pmax=(10.0, 'atm'), "
     "coeffs=[[1.0]])\n", "chebyshev_reaction unsupported"),
    ("ideal_gas(name='x', species='H' reactions='all')\n", "expected ',' or ')'"),
    ("ideal_gas(name='x', species='H', reactions='all'\n", "expected"),
])
def test_cti_errors_are_reported(tmp_path, text, err):
    f = tmp_path / "bad.cti"
    f.write_text(text)
    with pytest.raises(H.HpoError, match=re.escape(err)):
        H.Mechanism(str(f))


def test_cti_and_yaml_plog_read_alike(tmp_path):
    """A pressure-dependent Arrhenius reaction (CTI pdep_arrhenius, YAML
    pressure-dependent-Arrhenius) reads the same from both forms: the same kind, and the
    lowest-pressure level's SI-kmol parameters in the plain Arrhenius columns (the levels drive
    the kernels; tests/test_gpu_estimate.py checks them on ESTiMatE-Mech's PLOG rates)."""
    thermo = ("thermo=(NASA([200.0, 1000.0], [2.5, 0.0, 0.0, 0.0, 0.0, 25473.0, -0.44]), "
              "NASA([1000.0, 3500.0], [2.5, 0.0, 0.0, 0.0, 0.0, 25473.0, -0.44]))")
    cti = tmp_path / "p.cti"
    cti.write_text(f"""units(length='cm', time='s', quantity='mol', act_energy='cal/mol')
ideal_gas(name='p', elements='H', species='H H2', reactions='all')
species(name='H', atoms='H:1', {thermo})
species(name='H2', atoms='H:2', {thermo})
pdep_arrhenius('H + H <=> H2', [(10.0, 'atm'), 3.0e+13, 0.2, 1200.0], [(0.1, 'atm'), 1.0e+13, 0.0, 1000.0],
               [(1.0, 'atm'), 2.0e+13, 0.1, 1100.0])
""")
    nasa = "[2.5, 0.0, 0.0, 0.0, 0.0, 25473.0, -0.44]"
    yml = tmp_path / "p.yaml"
    yml.write_text(f"""units: {{length: cm, time: s, quantity: mol, activation-energy: cal/mol}}
phases:
- name: p
  thermo: ideal-gas
  elements: [H]
  species: [H, H2]
species:
- name: H
  composition: {{H: 1}}
  thermo: {{model: NASA7, temperature-ranges: [200.0, 1000.0, 3500.0], data: [{nasa}, {nasa}]}}
- name: H2
  composition: {{H: 2}}
  thermo: {{model: NASA7, temperature-ranges: [200.0, 1000.0, 3500.0], data: [{nasa}, {nasa}]}}
reactions:
- equation: H + H <=> H2
  type: pressure-dependent-Arrhenius
  rate-constants:
  - {{P: 0.1 atm, A: 1.0e+13, b: 0.0, Ea: 1000.0}}
  - {{P: 1.0 atm, A: 2.0e+13, b: 0.1, Ea: 1100.0}}
  - {{P: 10.0 atm, A: 3.0e+13, b: 0.2, Ea: 1200.0}}
""")
    c, y = H.Mechanism(str(cti)), H.Mechanism(str(yml))
    tc, ty = c.tables(), y.tables()
    assert list(tc["kind"]) == list(ty["kind"]) == [0]
    assert np.array_equal(tc["arr"], ty["arr"])
    assert tc["arr"][0, 3] == pytest.approx(1.0e13 * 1e-3, rel=1e-15)
    assert tc["arr"][0, 2] == pytest.approx(1000.0 * 4184.0 / 8314.4621, rel=1e-15)
