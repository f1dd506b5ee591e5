"""Gmsh 2.2 ASCII meshes (mesh_file *.msh), on the host (hifr_host_only solvers).

The *_gmsh_* fixtures come from the reference binary reading the same .msh files
through its own Gmsh reader (src/geometry.cpp:1309-1530, 1594-1640, 1881-2088);
test_gpu_parity.py steps the device from the .msh decks against them.  Here the
set-up from a .msh file must equal the set-up from the same mesh given as arrays
(element order, shape-node order, boundary faces, interfaces, metrics)."""
import os

import numpy as np
import pytest

import cases
import frmesh
import hifr


def gmsh_goldens():
    return [n for n in cases.golden_names() if str(cases.load_golden(n).get("mesh_format", "neu")) == "msh"]


def write_case(tmp_path, g, node_order=None, groups=None):
    ele_type, xv, c2v = cases.mesh_of(g)
    bf = cases.bfaces_of(g)
    walls = [] if bf is None else [(int(e), int(f)) for e, f, _ in bf]
    if groups is None:
        groups = [("CYCLIC", [(e, f) for e, f in frmesh.boundary_faces(ele_type, c2v) if (e, f) not in walls])]
        if walls:
            groups.append(("isotherm_fix", walls))
    frmesh.write_gmsh(str(tmp_path / "mesh.msh"), xv, c2v, ele_type, bc_groups=groups, node_order=node_order)
    deck = cases.product_deck(g, host_only=True, mesh_file="mesh.msh")
    (tmp_path / "input").write_text(deck)
    return str(tmp_path / "input")


def test_gmsh_goldens_cover_each_element_type():
    assert {int(cases.load_golden(n)["ele_type"]) for n in gmsh_goldens()} == {1, 2, 4}


@pytest.mark.parametrize("shuffle", [False, True], ids=["ordered", "shuffled_nodes"])
@pytest.mark.parametrize("name", gmsh_goldens())
def test_gmsh_setup_equals_array_setup(tmp_path, name, shuffle):
    g = cases.load_golden(name)
    order = np.random.default_rng(3).permutation(len(g["xv"])) if shuffle else None
    s = hifr.Solver.from_deck(write_case(tmp_path, g, node_order=order))
    a = cases.solver_of(g, hifr_host_only=1)
    assert (s.n_eles, s.nu, s.nfp, s.n_inters) == (a.n_eles, a.nu, a.nfp, a.n_inters)
    for what in ("xv", "c2v", "pos_upts", "detjac_upts", "JGinv_upts", "tdA_fpts", "norm_fpts"):
        assert np.array_equal(s.get(what), a.get(what)), what
    for x, y in zip(s.interfaces(), a.interfaces()):
        assert np.array_equal(x, y)


def test_gmsh_reader_errors(tmp_path):
    g = cases.load_golden("hex_p3_ns_channel_gmsh_3x2x2")
    ele_type, xv, c2v = cases.mesh_of(g)
    # a boundary face without a recognised condition name
    path = write_case(tmp_path, g, groups=[("no_such_bc", [(0, 0)])])
    with pytest.raises(RuntimeError, match="not recognized"):
        hifr.Solver.from_deck(path)
    # a mesh whose walls are not listed leaves faces without a neighbour
    path = write_case(tmp_path, g, groups=[("CYCLIC", [])])
    with pytest.raises(RuntimeError, match="no neighbour"):
        hifr.Solver.from_deck(path)
    # second-order elements are not read
    text = (tmp_path / "mesh.msh").read_text().replace("\n1 3 2 ", "\n1 10 2 ", 1)
    (tmp_path / "mesh.msh").write_text(text)
    with pytest.raises(RuntimeError):
        hifr.Solver.from_deck(path)
    # a binary or 4.x file is refused
    (tmp_path / "mesh.msh").write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
    with pytest.raises(RuntimeError, match="2.x"):
        hifr.Solver.from_deck(path)
