"""SBML -> input-folder converter on the CPU."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _read(d, name, dtype=float):
    return np.loadtxt(os.path.join(d, name), dtype=dtype, delimiter="\t", ndmin=2)


def test_schlogl_sbml_matches_golden_folder(tmp_path, ctl):
    from cutauleaping_amd.sbml import SBMLModel
    m = SBMLModel(os.path.join(GOLDEN, "schlogl_sbml.xml"))
    assert m.names == ["A_in_cell", "X_in_cell", "B_in_cell"]
    out = str(tmp_path / "f")
    m.write_folder(out, rows=1024, t_end=10.0, samples=100)
    g = os.path.join(GOLDEN, "schlogl")
    for name in ("left_side", "right_side", "M_feed"):
        assert np.array_equal(_read(out, name, int), _read(g, name, int)), name
    assert np.array_equal(_read(out, "MX_0"), _read(g, "MX_0"))
    assert np.allclose(_read(out, "c_matrix"), _read(g, "c_matrix"), rtol=0, atol=0)
    assert np.allclose(np.loadtxt(os.path.join(out, "t_vector")), np.loadtxt(os.path.join(g, "t_vector")))
    # the folder loads through the C ABI like the golden one
    a = ctl.Model.from_folder(out, 0, 1024).structure()
    b = ctl.Model.from_folder(g, 0, 1024).structure()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_reversible_global_parameters_add_reverse_row(tmp_path):
    from cutauleaping_amd.sbml import SBMLModel
    xml = open(os.path.join(GOLDEN, "schlogl_sbml.xml")).read()
    xml = xml.replace("</listOfReactions>", """<reaction id="R5" reversible="true">
        <listOfReactants><speciesReference species="X"/></listOfReactants>
        <listOfProducts><speciesReference species="A"/></listOfProducts>
        <kineticLaw><math xmlns="http://www.w3.org/1998/Math/MathML"><apply><minus/>
          <apply><times/><ci>kf</ci><ci>X</ci></apply><apply><times/><ci>kr</ci><ci>A</ci></apply></apply></math>
        </kineticLaw></reaction></listOfReactions>""")
    xml = xml.replace("<listOfReactions>", """<listOfParameters><parameter id="kf" value="0.5"/>
        <parameter id="kr" value="0.25"/></listOfParameters><listOfReactions>""")
    p = tmp_path / "m.xml"
    p.write_text(xml)
    m = SBMLModel(str(p))
    assert m.reactions[-2:] == ["R5", "R5 (reverse)"]
    assert m.left[-2] == [0, 1, 0] and m.right[-2] == [1, 0, 0]
    assert m.left[-1] == [1, 0, 0] and m.right[-1] == [0, 1, 0]
    assert m.c[-2:] == [0.5, 0.25]


def test_reference_schlogl_xml_when_present(tmp_path):
    ref = "/root/reference/schlogl.xml"
    if not os.path.exists(ref):
        pytest.skip("reference checkout not mounted")
    from cutauleaping_amd.sbml import SBMLModel
    m = SBMLModel(ref)
    assert [n.split("_in_")[0] for n in m.names] == ["A", "X", "B"]
    assert m.left == [[1, 2, 0], [0, 3, 0], [0, 0, 1], [0, 1, 0]]
    assert m.right == [[0, 3, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1]]
    assert m.c == [3e-7, 1e-4, 1e-3, 3.5]
    assert m.x0 == [100000.0, 250.0, 200000.0] and m.feed == [1, 0, 1]
