"""MaHyCo case files (mahyco_amd.arc): the reference's ctest input
tests/golden/Data.1.arc (mahyco/data/Data.1.arc, a 250x25x25 two-environment
Sod along x) read into mhg_options, checked value by value against the file,
and run on the oracle at a reduced mesh; the older layout of
mahyco/src/Mahyco.arc (cas-test and ordre-projection directly under
<mahyco>); options the library does not run are refused by name."""
import os

import numpy as np
import pytest

import helpers as H
from mahyco_amd import abi, arc

HERE = os.path.dirname(os.path.abspath(__file__))
DATA1 = os.path.join(HERE, "golden", "Data.1.arc")


def test_data1_options():
    c = arc.read_case(DATA1)
    assert c.dim == 3 and c.n == [250, 25, 25]
    assert c.origin == [0.0, 0.0, 0.0] and c.length == [1.0, 0.1, 0.1]
    assert [(e[0], e[1], e[2]) for e in c.envs] == [("ZG", abi.EOS_PERFECT_GAS, 1.4), ("ZD", abi.EOS_PERFECT_GAS, 1.4)]
    assert c.cas_model == "SOD" and c.cas_test == 13
    o = c.options
    assert o["ordre_projection"] == 2 and o["pseudo_centree"] == 0 and o["schema_csts"] == 0
    assert o["deltat_init"] == 1e-5 and o["deltat_min"] == 1e-8 and o["deltat_max"] == 0.01
    assert o["final_time"] == 0.2 and o["longueur_caracteristique"] == abi.LC_RACINE_CUBIQUE_VOLUME
    assert [b[0] for b in c.bcs] == ["XMIN", "XMAX", "YMIN", "YMAX", "ZMIN", "ZMAX"]
    assert [b[1] for b in c.bcs] == [abi.BC_VX, abi.BC_VX, abi.BC_VY, abi.BC_VY, abi.BC_VZ, abi.BC_VZ]
    assert c.ignored == []


def test_data1_runs_on_the_oracle():
    """At 40x3x3 cells the case runs as the hand-built 3D two-environment Sod
    with the same options does, bit for bit."""
    c = arc.read_case(DATA1)
    L = H.oracle()
    s = c.build(L, n=[40, 3, 3])
    s.run(30)
    a = s.download()
    opt = L.options(deltat_init=1e-5, deltat_min=1e-8, deltat_max=0.01, final_time=0.2)
    abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, 1.4, 0.), (abi.EOS_PERFECT_GAS, 1.4, 0.)])
    abi.set_bcs(opt, abi.reflecting_walls(3))
    r = abi.Simulation(L, abi.mesh_desc(3, [40, 3, 3], [0, 0, 0], [1.0, 0.1, 0.1]), opt)
    r.init_sod(13)
    r.run(30)
    b = r.download()
    for f in ["density", "pressure", "velocity", "fracvol_env"]:
        assert np.array_equal(a[f], b[f]), f
    assert a["global_iteration"] == 30 and a["global_time"] > 0
    s.close()
    r.close()


LEGACY = """<?xml version='1.0'?>
<case codeversion="1.0" codename="Mahyco" xml:lang="en">
  <mesh><meshgenerator><cartesian><origine>0.0 0.0</origine>
    <lx nx='20' prx='1.0'>1.</lx><ly ny='2' pry='1.0'>.1</ly></cartesian></meshgenerator></mesh>
  <mahyco>
    <material><name>ZG_mat</name></material>
    <environment><name>ZG</name><material>ZG_mat</material>
      <eos-model name="StiffenedGas"><adiabatic-cst>1.6</adiabatic-cst><limit-tension>0.01</limit-tension></eos-model>
    </environment>
    <cas-test>3</cas-test>
    <ordre-projection>1</ordre-projection>
    <projection-limiteur-id>5</projection-limiteur-id>
    <deltat-init>1e-5</deltat-init><deltat-min>1e-8</deltat-min><deltat-max>1e-3</deltat-max>
    <final-time>0.1</final-time><cfl>0.25</cfl>
    <longueur-caracteristique>faces-opposees</longueur-caracteristique>
    <convnerg>false</convnerg>
    <boundary-condition><surface>ZMIN</surface><type>Vz</type><value>0.</value></boundary-condition>
    <boundary-condition><surface>XMIN</surface><type>Vx</type><value>0.</value></boundary-condition>
  </mahyco>
</case>"""


def test_older_layout_and_2d():
    c = arc.read_case(LEGACY)
    assert c.dim == 2 and c.n == [20, 2] and c.length == [1.0, 0.1]
    assert c.cas_test == 3 and c.options["ordre_projection"] == 1
    assert c.options["projection_limiteur_id"] == abi.LIM_VANLEERG and c.options["cfl"] == 0.25
    assert c.options["longueur_caracteristique"] == abi.LC_FACES_OPPOSEES
    assert c.envs == [("ZG", abi.EOS_STIFFENED_GAS, 1.6, 0.01)]
    assert c.bcs == [("XMIN", abi.BC_VX, 0.0)]  # no z faces in 2D
    assert c.ignored == ["convnerg"]
    s = c.build(H.oracle())
    s.run(5)
    s.close()


@pytest.mark.parametrize("change,what", [
    ("<ordre-projection>1</ordre-projection>", "<calcul_flux_masse>1</calcul_flux_masse>"),
    ("prx='1.0'", "prx='1.1'"),
    ("StiffenedGas", "MieGruneisen"),
    ("<cas-test>3</cas-test>", "<cas-test>7</cas-test>"),
])
def test_unsupported_options_refused(change, what):
    with pytest.raises(arc.UnsupportedOption):
        arc.read_case(LEGACY.replace(change, what))


def test_missing_required_option():
    with pytest.raises(ValueError, match="deltat-max"):
        arc.read_case(LEGACY.replace("<deltat-max>1e-3</deltat-max>", ""))
