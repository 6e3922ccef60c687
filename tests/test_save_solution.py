"""HDF5 solution files (libtrixih5.so, include/trixi_h5.h).

The layout follows Trixi.jl's save_solution_file
(src/callbacks_step/save_solution_dg.jl), which the reference's examples use
with solution_variables = cons2prim (examples/euler_ec_1d.jl:37-39).  Files
are read back with the HDF5 tools' h5dump: raw little-endian dumps of each
dataset (compared bit for bit) and the attribute listing.  cons2prim is
pinned by round trips through prim2cons (an independent scalar
implementation) and, for ideal MHD, by a state whose pressure is worked out
by hand.
"""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
needs_h5 = pytest.mark.skipif(
    not os.path.exists(os.path.join(ROOT, "trixicuda.jl_amd", "libtrixih5.so")) or H5DUMP is None,
    reason="libtrixih5.so or h5dump absent (no HDF5 library)")


def h5_dataset(path, name, tmp_path):
    out = str(tmp_path / (name.strip("/") + ".bin"))
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, capture_output=True)
    return np.fromfile(out, dtype="<f8")


def h5_attributes(path):
    """{attribute path: text value} from h5dump's listing of the header and attributes."""
    text = subprocess.run([H5DUMP, "-A", path], check=True, capture_output=True, text=True).stdout
    attrs, where = {}, []
    for m in re.finditer(r'(GROUP|DATASET|ATTRIBUTE) "([^"]+)" \{|DATA \{\s*\(0\): ([^\n]+)', text):
        if m.group(1) in ("GROUP", "DATASET"):
            where = [m.group(2)]
        elif m.group(1) == "ATTRIBUTE":
            attrs[(where[0] if where else "/", m.group(2))] = None
            last = (where[0] if where else "/", m.group(2))
        else:
            attrs[last] = m.group(3).strip().strip('"')
    return attrs, text


def test_header_declares_the_writer():
    with open(os.path.join(ROOT, "include", "trixi_h5.h")) as fh:
        text = fh.read()
    syms = set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(trixi_h5_\w+)\s*\(", text, re.M))
    assert syms == {"trixi_h5_save_solution", "trixi_h5_save_mesh", "trixi_h5_last_error"}
    if os.path.exists(os.path.join(ROOT, "trixicuda.jl_amd", "libtrixih5.so")):
        lib = ctypes.CDLL(os.path.join(ROOT, "trixicuda.jl_amd", "libtrixih5.so"))
        for s in syms:
            assert hasattr(lib, s), s


@needs_h5
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_written_file_matches_trixi_layout(th, tmp_path, nd):
    rng = np.random.default_rng(nd)
    eq = [th.CompressibleEulerEquations1D, th.CompressibleEulerEquations2D, th.CompressibleEulerEquations3D][nd - 1](1.4)
    E, N = 5, 4
    data = rng.standard_normal((E,) + (N,) * nd + (nd + 2,))
    alpha = rng.random(E)
    path = str(tmp_path / "solution_000000007.h5")
    th.write_solution_file(path, data, eq, 3, t=0.25, dt=0.0125, timestep=7, solution_variables=th.cons2prim,
                           element_variables={"indicator_shock_capturing": alpha})
    for v in range(nd + 2):  # variables_v: Trixi's u[v, i, j, k, e] column-major, i fastest
        got = h5_dataset(path, f"/variables_{v + 1}", tmp_path)
        assert got.size == E * N ** nd
        np.testing.assert_array_equal(got, data[..., v].reshape(-1))
        # element e, node (i, j, k) sits at i + N j + N^2 k + N^nd e
        e, i, j, k = 2, 3, 1, 2
        jk = (j, i) if nd == 2 else (k, j, i) if nd == 3 else (i,)
        off = i + (N * j if nd > 1 else 0) + (N * N * k if nd > 2 else 0) + N ** nd * e
        assert got[off] == data[(e,) + jk + (v,)]
    np.testing.assert_array_equal(h5_dataset(path, "/element_variables_1", tmp_path), alpha)
    attrs, text = h5_attributes(path)
    expect = {"ndims": str(nd), "equations": f"CompressibleEulerEquations{nd}D", "polydeg": "3",
              "n_vars": str(nd + 2), "n_elements": str(E), "mesh_type": "TreeMesh", "mesh_file": "",
              "timestep": "7"}
    for k, v in expect.items():
        assert attrs[("/", k)] == v, (k, attrs.get(("/", k)))
    assert float(attrs[("/", "time")]) == 0.25 and float(attrs[("/", "dt")]) == 0.0125
    names = ["rho"] + [f"v{d + 1}" for d in range(nd)] + ["p"]
    for v, n in enumerate(names):
        assert attrs[(f"variables_{v + 1}", "name")] == n
    assert attrs[("element_variables_1", "name")] == "indicator_shock_capturing"
    assert "H5T_STD_I64LE" in text and "H5T_IEEE_F64LE" in text and "H5T_CSET_UTF8" in text


@needs_h5
def test_writer_rejections(th, tmp_path):
    eq = th.CompressibleEulerEquations2D(1.4)
    with pytest.raises(ValueError, match="shape"):
        th.write_solution_file(str(tmp_path / "a.h5"), np.zeros((3, 4, 5, 4)), eq, 3)
    with pytest.raises(ValueError, match="variables"):
        th.write_solution_file(str(tmp_path / "a.h5"), np.zeros((3, 4, 4, 5)), eq, 3)
    with pytest.raises(ValueError, match="one value per element"):
        th.write_solution_file(str(tmp_path / "a.h5"), np.zeros((3, 4, 4, 4)), eq, 3,
                               element_variables={"x": np.zeros(2)})
    with pytest.raises(th.TrixiHipError, match="cannot create"):
        th.write_solution_file(str(tmp_path / "missing_dir" / "a.h5"), np.zeros((3, 4, 4, 4)), eq, 3)
    with pytest.raises(ValueError, match="cons2prim or cons2cons"):
        th.SaveSolutionCallback(solution_variables=lambda u, e: u)
    with pytest.raises(ValueError, match="interval or a positive dt"):
        th.SaveSolutionCallback(interval=10, dt=0.1)
    with pytest.raises(ValueError, match="interval or a positive dt"):
        th.SaveSolutionCallback(dt=0.0)
    with pytest.raises(RuntimeError, match="initialize"):
        th.SaveSolutionCallback(dt=0.1)(None, 0.1, 0.1, 1, None)


def _families(th):
    yield th.CompressibleEulerEquations1D(1.4), [1.3, -0.4, 2.1]
    yield th.CompressibleEulerEquations2D(1.4), [0.9, 0.3, -0.7, 1.6]
    yield th.CompressibleEulerEquations3D(1.67), [1.1, 0.2, 0.5, -0.3, 0.8]
    yield th.ShallowWaterEquations1D(gravity_constant=9.81), [3.0, 0.4, 0.7]
    yield th.ShallowWaterEquations2D(gravity_constant=9.81, H0=3.0), [3.0, 0.4, -0.2, 0.7]
    yield th.CompressibleEulerEquationsQuasi1D(1.4), [1.2, 0.5, 1.9, 0.6]
    yield th.CompressibleEulerMulticomponentEquations1D(gammas=(1.4, 1.648), gas_constants=(0.4, 2.4)), \
        [0.3, 1.7, 0.8, 0.1]
    yield th.CompressibleEulerMulticomponentEquations2D(gammas=(1.4, 1.648, 1.3), gas_constants=(0.4, 2.4, 1.1)), \
        [0.3, -0.2, 1.7, 0.8, 0.1, 0.5]


def test_cons2prim_inverts_prim2cons(th):
    for eq, prim in _families(th):
        u = np.array(th.prim2cons(prim, eq))
        np.testing.assert_allclose(th.cons2prim(u, eq), prim, rtol=1e-13, atol=1e-15, err_msg=type(eq).__name__)
        batch = np.stack([u, u * 1.0])[None]  # vectorized over leading axes
        np.testing.assert_allclose(th.cons2prim(batch, eq)[0, 1], prim, rtol=1e-13)
        assert len(th.varnames(th.cons2prim, eq)) == len(th.varnames(th.cons2cons, eq)) == eq.nvariables
        np.testing.assert_array_equal(th.cons2cons(u, eq), u)


def test_cons2prim_mhd_hand_state(th):
    # rho 2, v (1, 0, 0), p 1, B (1, 0, 0), psi 0.5, gamma 5/3:
    # rho e = p/(gamma-1) + rho |v|^2/2 + |B|^2/2 + psi^2/2 = 1.5 + 1 + 0.5 + 0.125
    eq3 = th.IdealGlmMhdEquations3D(5.0 / 3.0)
    u = np.array([2.0, 2.0, 0.0, 0.0, 3.125, 1.0, 0.0, 0.0, 0.5])
    np.testing.assert_allclose(th.cons2prim(u, eq3), [2.0, 1.0, 0, 0, 1.0, 1.0, 0, 0, 0.5], rtol=1e-15)
    eq1 = th.IdealGlmMhdEquations1D(5.0 / 3.0)  # no psi: rho e = 3.0
    np.testing.assert_allclose(th.cons2prim(u[:8] - np.eye(8)[4] * 0.125, eq1), [2.0, 1.0, 0, 0, 1.0, 1.0, 0, 0],
                               rtol=1e-15)
    assert th.varnames(th.cons2prim, eq3) == ["rho", "v1", "v2", "v3", "p", "B1", "B2", "B3", "psi"]
    assert th.varnames(th.cons2cons, eq1) == ["rho", "rho_v1", "rho_v2", "rho_v3", "rho_e", "B1", "B2", "B3"]


def test_varnames_of_every_family(th):
    assert th.varnames(th.cons2prim, th.LinearScalarAdvectionEquation2D((1.0, 1.0))) == ["scalar"]
    assert th.varnames(th.cons2cons, th.InviscidBurgersEquation1D()) == ["scalar"]
    assert th.varnames(th.cons2prim, th.HyperbolicDiffusionEquations3D()) == ["phi", "q1", "q2", "q3"]
    assert th.varnames(th.cons2cons, th.CompressibleEulerEquations2D(1.4)) == ["rho", "rho_v1", "rho_v2", "rho_e"]
    m = th.CompressibleEulerMulticomponentEquations2D(gammas=(1.4, 1.648), gas_constants=(0.4, 2.4))
    assert th.varnames(th.cons2cons, m) == ["rho_v1", "rho_v2", "rho_e", "rho1", "rho2"]
    assert th.varnames(th.cons2prim, m) == ["v1", "v2", "p", "rho1", "rho2"]
    q = th.CompressibleEulerEquationsQuasi1D(1.4)
    assert th.varnames(th.cons2cons, q) == ["a_rho", "a_rho_v1", "a_e", "a"]
    s1 = th.ShallowWaterEquations1D(gravity_constant=9.81)
    assert th.varnames(th.cons2cons, s1) == ["h", "h_v", "b"] and th.varnames(th.cons2prim, s1) == ["H", "v", "b"]
    s2 = th.ShallowWaterEquations2D(gravity_constant=9.81)
    assert th.varnames(th.cons2prim, s2) == ["H", "v1", "v2", "b"]


def h5_int_dataset(path, name, tmp_path):
    out = str(tmp_path / (name.strip("/") + ".bin"))
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], check=True, capture_output=True)
    return np.fromfile(out, dtype="<i8")


# the reference's mesh shapes: uniform periodic (test/tree_dgsem_3d/euler_ec.jl),
# non-periodic, and refined with the same box applied twice
# (test/tree_dgsem_2d/euler_blob_mortar.jl:59-68)
MESHES = [
    dict(lo=(-2.0,) * 3, hi=(2.0,) * 3, level=2, periodicity=True, patches=()),
    dict(lo=(0.0, 0.0), hi=(2.0, 2.0), level=3, periodicity=(False, True), patches=()),
    dict(lo=(-32.0, -32.0), hi=(32.0, 32.0), level=4, periodicity=True,
         patches=[dict(type="box", coordinates_min=(-40.0, -5.0), coordinates_max=(40.0, 5.0))] * 2),
    dict(lo=(-1.0,), hi=(1.0,), level=4, periodicity=False, patches=()),
]


@pytest.mark.parametrize("spec", MESHES, ids=["uniform3d", "nonperiodic2d", "blob2d", "uniform1d"])
def test_tree_arrays_are_a_consistent_treemesh(th, spec):
    """The cell tree written to mesh.h5: the leaves in storage order are the
    elements (same centres and levels), storage order is depth first (each
    parent directly before its first child, children consecutive), a child's
    centre is its parent's shifted by a quarter of the parent's length along
    each axis, and neighbour links are symmetric and join same-level cells
    whose centres differ by one cell length along one axis (modulo the
    domain length in periodic directions)."""
    m = th.TreeMesh(spec["lo"], spec["hi"], spec["level"], periodicity=spec["periodicity"],
                    refinement_patches=spec["patches"])
    t = m.tree_arrays()
    nd, n = m.ndims, t["n_cells"]
    P, C, Nb, lev, X = t["parent_ids"], t["child_ids"], t["neighbor_ids"], t["levels"], t["coordinates"]
    leaf = np.all(C == 0, axis=1)
    assert t["n_leaf_cells"] == m.nelements == leaf.sum()
    np.testing.assert_array_equal(X[leaf], m.centers)
    assert P[0] == 0 and lev[0] == 0
    np.testing.assert_array_equal(X[0], 0.5 * (np.array(spec["lo"]) + np.array(spec["hi"])))
    L0 = t["length_level_0"]
    for i in range(n):
        kids = C[i][C[i] > 0]
        if kids.size:
            assert kids.size == 1 << nd and kids[0] == i + 2 and np.all(np.diff(kids) > 0)
            for c, k in enumerate(C[i]):
                assert P[k - 1] == i + 1 and lev[k - 1] == lev[i] + 1
                off = np.array([0.5 if (c >> d) & 1 else -0.5 for d in range(nd)]) * L0 / 2 ** (lev[i] + 1)
                np.testing.assert_array_equal(X[k - 1], X[i] + off)
        per = m.periodicity
        for f in range(2 * nd):
            j = Nb[i, f]
            if j == 0:
                continue
            d, side = f // 2, f % 2
            assert lev[j - 1] == lev[i] and Nb[j - 1, f ^ 1] == i + 1
            h = L0 / 2 ** lev[i]
            delta = X[j - 1] - X[i]
            expect = np.zeros(nd)
            expect[d] = h if side else -h
            diff = delta - expect
            if per[d]:
                diff[d] = (diff[d] + L0 / 2) % L0 - L0 / 2 if lev[i] > 0 else 0.0
            np.testing.assert_allclose(diff, 0.0, atol=1e-12 * L0)
    # a same-level neighbour exists exactly where a cell's neighbour position is inside the tree
    if not any(m.periodicity):
        root_nbrs = Nb[0]
        assert np.all(root_nbrs == 0)


@needs_h5
def test_mesh_file_matches_the_tree(th, tmp_path):
    """mesh.h5 as Trixi's save_mesh_file writes it for a TreeMesh: the tree
    datasets (row-major transposes of Trixi's column-major arrays) and the
    root attributes, read back with h5dump."""
    m = th.TreeMesh((-32.0, -32.0), (32.0, 32.0), 4, n_cells_max=100_000,
                    refinement_patches=[dict(type="box", coordinates_min=(-40.0, -5.0),
                                             coordinates_max=(40.0, 5.0))] * 2)
    path = th.save_mesh_file(m, str(tmp_path))
    assert os.path.basename(path) == "mesh.h5"
    t = m.tree_arrays()
    for k in ("parent_ids", "child_ids", "neighbor_ids", "levels"):
        np.testing.assert_array_equal(h5_int_dataset(path, "/" + k, tmp_path), t[k].reshape(-1))
    np.testing.assert_array_equal(h5_dataset(path, "/coordinates", tmp_path), t["coordinates"].reshape(-1))
    attrs, text = h5_attributes(path)
    assert attrs[("/", "mesh_type")] == "TreeMesh"
    assert int(attrs[("/", "ndims")]) == 2
    assert int(attrs[("/", "n_cells")]) == t["n_cells"]
    assert int(attrs[("/", "capacity")]) == 100_000
    assert int(attrs[("/", "n_leaf_cells")]) == m.nelements
    assert int(attrs[("/", "minimum_level")]) == 4 and int(attrs[("/", "maximum_level")]) == 6
    assert float(attrs[("/", "length_level_0")]) == 64.0
    for a in ("center_level_0", "periodicity"):
        assert ("/", a) in attrs
    assert re.search(r'DATASET "child_ids" \{\s*DATATYPE\s+H5T_STD_I64LE\s*DATASPACE\s+SIMPLE \{ \( %d, 4 \)'
                     % t["n_cells"], subprocess.run([H5DUMP, "-H", path], check=True, capture_output=True,
                                                    text=True).stdout)
