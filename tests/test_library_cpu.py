"""CPU-side checks of the product: the C-ABI library loads and exports every
symbol include/trixi_hip.h declares; its host-side basis matches the exact
fixtures; the host TreeMesh matches the oracle's; solver configuration maps
to the right enums and rejects unsupported combinations.  No device call."""
import json
import os
import re

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    with open(os.path.join(ROOT, "include", "trixi_hip.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(trixi_hip_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    for s in ("trixi_hip_create", "trixi_hip_rhs", "trixi_hip_rk_step", "trixi_hip_max_dt",
              "trixi_hip_error_norms", "trixi_hip_basis", "trixi_hip_pack", "trixi_hip_unpack"):
        assert s in syms


def test_library_exports_every_declared_symbol(th):
    import ctypes
    L = ctypes.CDLL(th.LIBRARY_PATH)
    for s in declared_symbols():
        assert hasattr(L, s), s
    from trixi_hip import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES), "ctypes binding out of sync with the header"
    assert "gfx950" in th.version()


@pytest.mark.parametrize("N", [2, 3, 4, 5, 6])
def test_library_basis_matches_exact_fixture(th, N):
    with open(os.path.join(GOLDEN, f"basis_N{N}.json")) as fh:
        g = {k: np.array(v, dtype=np.float64) for k, v in json.load(fh).items() if k != "N"}
    b = th.LobattoLegendreBasisHIP(N - 1)
    np.testing.assert_allclose(b.nodes, g["nodes"], rtol=0, atol=2e-16 * 4)
    np.testing.assert_allclose(b.weights, g["weights"], rtol=0, atol=4e-16)
    for key, mine in (("D", b.derivative_matrix), ("Dhat", b.derivative_dhat), ("Vinv", b.inverse_vandermonde_legendre)):
        np.testing.assert_allclose(mine, g[key], rtol=0, atol=2e-14 * np.abs(g[key]).max(), err_msg=key)
    np.testing.assert_allclose(b.derivative_split, g["Dsplit"], rtol=0, atol=2e-14 * np.abs(g["Dsplit"]).max())
    np.testing.assert_allclose(b.boundary_interpolation[:, 0], g["lhat_minus"], atol=1e-14)
    np.testing.assert_allclose(b.boundary_interpolation[:, 1], g["lhat_plus"], atol=1e-14)


@pytest.mark.parametrize("nd,level", [(1, 4), (2, 3), (3, 2)])
def test_host_treemesh_matches_oracle(th, oracle, nd, level):
    m = th.TreeMesh((-2.0,) * nd, (2.0,) * nd, initial_refinement_level=level, n_cells_max=10**6)
    o = oracle.Mesh(nd, level, -2.0, 2.0, 3)
    assert m.nelements == o.nelements and m.ninterfaces == o.ninterfaces
    np.testing.assert_array_equal(m.neighbor_ids, o.neighbor_ids)
    np.testing.assert_array_equal(m.orientations, o.orientations)
    np.testing.assert_array_equal(m.cell_index, o.cell_index)
    np.testing.assert_array_equal(m.inverse_jacobian, o.inverse_jacobian)
    x = m.node_coordinates(th.LobattoLegendreBasisHIP(3).nodes)
    np.testing.assert_allclose(x, o.node_coordinates, rtol=0, atol=1e-15)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_nonperiodic_mesh_matches_oracle(th, oracle, nd):
    lvl = 2
    m = th.TreeMesh((0.0,) * nd, (2.0,) * nd, lvl, periodicity=False)
    om = oracle.Mesh(nd, lvl, 0.0, 2.0, 3, periodic=False)
    assert m.ninterfaces == om.ninterfaces == nd * (m.nelements - m.nelements // (1 << lvl))
    np.testing.assert_array_equal(m.neighbor_ids, om.neighbor_ids)
    np.testing.assert_array_equal(m.orientations, om.orientations)
    assert len(m.boundary_faces) == 2 * nd * (1 << (lvl * (nd - 1)))


def test_config_mapping_and_rejections(th):
    eq = th.CompressibleEulerEquations3D(1.4)
    s = th.DGSEMHIP(polydeg=3, surface_flux=th.flux_lax_friedrichs,
                    volume_integral=th.VolumeIntegralFluxDifferencing(th.flux_ranocha))
    c = s.config(eq)
    assert (c.ndims, c.equation, c.polydeg, c.volume_integral, c.volume_flux, c.surface_flux, c.precision) == \
        (3, 1, 3, 1, 3, 1, 0)
    s32 = th.DGSEMHIP(polydeg=3, surface_flux=th.FluxLaxFriedrichs(th.max_abs_speed_naive),
                      volume_integral=th.VolumeIntegralFluxDifferencing(th.flux_shima_etal), RealT=np.float32)
    c = s32.config(eq)
    # RealT = Float32 is the FP32-state precision (TRIXI_HIP_FP32 = 3)
    assert (c.surface_flux, c.volume_flux, c.precision) == (2, 4, 3)
    assert th.DGSEMHIP(polydeg=3, precision="mixed").config(eq).precision == 1
    mhd = th.IdealGlmMhdEquations3D(5 / 3)
    with pytest.raises(NotImplementedError):
        th.DGSEMHIP(polydeg=3, surface_flux=th.flux_lax_friedrichs).config(mhd)
    ok = th.DGSEMHIP(polydeg=3, surface_flux=(th.flux_lax_friedrichs, th.flux_nonconservative_powell),
                     volume_integral=th.VolumeIntegralFluxDifferencing(
                         (th.flux_hindenlang_gassner, th.flux_nonconservative_powell)))
    c = ok.config(mhd)
    assert (c.equation, c.volume_flux, c.surface_flux, c.c_h) == (2, 5, 1, 1.0)
    # mhd_shock.jl: HG with (flux, flux_nonconservative_powell) as DG and FV flux
    pair = (th.flux_hindenlang_gassner, th.flux_nonconservative_powell)
    ind_mhd = th.IndicatorHennemannGassner(mhd, th.LobattoLegendreBasisHIP(4), alpha_max=0.5, alpha_min=0.001,
                                           alpha_smooth=True, variable=th.density_pressure)
    hg = th.DGSEMHIP(polydeg=4, surface_flux=pair, volume_integral=th.VolumeIntegralShockCapturingHG(
        ind_mhd, volume_flux_dg=pair, volume_flux_fv=pair)).config(mhd)
    assert (hg.volume_integral, hg.volume_flux, hg.fv_flux, hg.surface_flux, hg.indicator_variable) == (2, 5, 5, 5, 0)
    with pytest.raises(NotImplementedError):
        th.DGSEMHIP(polydeg=4, surface_flux=pair, volume_integral=th.VolumeIntegralShockCapturingHG(
            ind_mhd, volume_flux_dg=th.flux_hindenlang_gassner, volume_flux_fv=pair)).config(mhd)
    eq2 = th.CompressibleEulerEquations2D(1.4)
    b = th.LobattoLegendreBasisHIP(3)
    ind = th.IndicatorHennemannGassner(eq2, b, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True,
                                       variable=th.density_pressure)
    hg = th.DGSEMHIP(polydeg=3, surface_flux=th.flux_lax_friedrichs,
                     volume_integral=th.VolumeIntegralShockCapturingHG(ind, volume_flux_dg=th.flux_ranocha,
                                                                       volume_flux_fv=th.flux_lax_friedrichs))
    c = hg.config(eq2)
    assert (c.volume_integral, c.volume_flux, c.fv_flux, c.alpha_smooth, c.alpha_max) == (2, 3, 1, 1, 0.5)
    m = th.TreeMesh((-1.0,), (1.0,), 2, periodicity=False)
    assert m.ninterfaces == 3 and m.boundary_faces.tolist() == [[0, 0], [3, 1]]
    with pytest.raises(ValueError):  # a non-periodic mesh needs boundary conditions
        th.SemidiscretizationHyperbolicHIP(m, th.CompressibleEulerEquations1D(1.4),
                                           th.initial_condition_convergence_test,
                                           th.DGSEMHIP(polydeg=3, surface_flux=th.flux_lax_friedrichs))
    with pytest.raises(NotImplementedError):  # one Dirichlet state for all sides
        th.SemidiscretizationHyperbolicHIP(
            m, th.CompressibleEulerEquations1D(1.4), th.initial_condition_convergence_test,
            th.DGSEMHIP(polydeg=3, surface_flux=th.flux_lax_friedrichs),
            boundary_conditions={"x_neg": th.BoundaryConditionDirichlet(th.initial_condition_convergence_test),
                                 "x_pos": th.BoundaryConditionDirichlet(th.initial_condition_constant)})
    with pytest.raises(ValueError):
        th.TreeMesh((-1.0, -1.0), (1.0, 2.0), 2)


def test_device_calls_fail_loudly_without_gpu(th):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    mesh = th.TreeMesh((-1.0,), (1.0,), 2)
    solver = th.DGSEMHIP(polydeg=3, surface_flux=th.flux_lax_friedrichs)
    with pytest.raises(th.TrixiHipError):
        th.SemidiscretizationHyperbolicHIP(mesh, th.LinearScalarAdvectionEquation1D(1.0),
                                           th.initial_condition_convergence_test, solver)


# the reference's mortar meshes: test/tree_dgsem_2d/advection_mortar.jl:16-23,
# test/tree_dgsem_3d/advection_mortar.jl:16-25, test/tree_dgsem_3d/euler_mortar.jl:15-22,
# test/tree_dgsem_2d/euler_blob_mortar.jl:59-68
REFERENCE_MORTAR_MESHES = [
    (2, 2, -1.0, 1.0, [((0.0, -1.0), (1.0, 1.0))]),
    (3, 2, -1.0, 1.0, [((0.0, -1.0, -1.0), (1.0, 1.0, 1.0)), ((0.0, -0.5, -0.5), (0.5, 0.5, 0.5))]),
    (3, 2, 0.0, 2.0, [((0.5,) * 3, (1.5,) * 3)]),
    (2, 4, -32.0, 32.0, [((-40.0, -5.0), (40.0, 5.0)), ((-40.0, -5.0), (40.0, 5.0))]),
]


@pytest.mark.parametrize("spec", REFERENCE_MORTAR_MESHES, ids=["adv2d", "adv3d", "euler3d", "blob2d"])
@pytest.mark.parametrize("periodic", [True, False])
def test_refined_treemesh_matches_oracle(th, oracle, spec, periodic):
    """The host tree (Trixi's refine!/rebalance! replayed cell by cell) and the
    oracle's geometric construction (minimal 2:1 balance, leaves in Morton
    order) give the same elements, interfaces, mortars and boundary faces."""
    nd, lvl, a, b, patches = spec
    m = th.TreeMesh((a,) * nd, (b,) * nd, lvl, periodicity=periodic,
                    refinement_patches=[dict(type="box", coordinates_min=lo, coordinates_max=hi) for lo, hi in patches])
    o = oracle.RefinedMesh(nd, lvl, a, b, 3, patches, periodic=periodic)
    assert m.nelements == o.nelements and m.ninterfaces == o.ninterfaces and m.nmortars == o.nmortars > 0
    np.testing.assert_array_equal(m.cell_index, o.cell_index)
    np.testing.assert_array_equal(m.levels, o.levels)
    np.testing.assert_array_equal(m.neighbor_ids, o.neighbor_ids)
    np.testing.assert_array_equal(m.orientations, o.orientations)
    np.testing.assert_array_equal(m.boundary_faces, o.boundary_faces)
    np.testing.assert_array_equal(m.inverse_jacobian, o.inverse_jacobian)
    key = lambda X: sorted(zip(map(tuple, X.mortar_neighbor_ids), X.mortar_orientations, X.mortar_large_sides))
    assert key(m) == key(o)
    x = m.node_coordinates(th.LobattoLegendreBasisHIP(3).nodes)
    np.testing.assert_array_equal(x, o.node_coordinates)


def test_mortar_mesh_rejections(th):
    with pytest.raises(NotImplementedError):
        th.TreeMesh((-1.0,), (1.0,), 2, refinement_patches=[dict(type="box", coordinates_min=(0.0,),
                                                                  coordinates_max=(1.0,))])
    m = th.TreeMesh((-1.0, -1.0), (1.0, 1.0), 2,
                    refinement_patches=[dict(type="box", coordinates_min=(0.0, -1.0), coordinates_max=(1.0, 1.0))])
    bad = m.mortar_neighbor_ids.copy()
    bad[0, 0] = m.nelements  # out of range: rejected before any device call
    m.mortar_neighbor_ids = bad
    with pytest.raises(th.TrixiHipError, match="mortar 0 out of range"):
        th.SemidiscretizationHyperbolicHIP(m, th.LinearScalarAdvectionEquation2D((0.2, -0.7)),
                                           th.initial_condition_convergence_test,
                                           th.DGSEMHIP(polydeg=3, surface_flux=th.flux_lax_friedrichs))


def test_prim2cons_and_quasi1d_flux_pairs(th):
    """prim2cons for the function-written initial conditions, and the
    (flux, nonconservative flux) pairs quasi-1D Euler requires."""
    e1 = th.CompressibleEulerEquations1D(1.4)
    assert th.prim2cons((2.0, 0.5, 1.0), e1) == pytest.approx([2.0, 1.0, 1.0 / 0.4 + 0.25])
    q = th.CompressibleEulerEquationsQuasi1D(1.4)
    assert q.nvariables == 4
    assert th.prim2cons((2.0, 0.5, 1.0, 3.0), q) == pytest.approx([6.0, 3.0, 3.0 * (2.5 + 0.25), 3.0])
    m = th.CompressibleEulerMulticomponentEquations2D(gammas=(1.4, 2.0), gas_constants=(0.4, 1.0))
    # gamma of the mixture: sum rho_k c_v,k gamma_k / sum rho_k c_v,k, c_v = R / (gamma - 1)
    gam = (1.0 * 1.0 * 1.4 + 3.0 * 1.0 * 2.0) / (1.0 * 1.0 + 3.0 * 1.0)
    assert th.prim2cons((0.5, 0.0, 1.0, 1.0, 3.0), m) == pytest.approx(
        [2.0, 0.0, 1.0 / (gam - 1) + 0.5 * 4.0 * 0.25, 1.0, 3.0])
    pair = (th.flux_chan_etal, th.flux_nonconservative_chan_etal)
    c = th.DGSEMHIP(polydeg=4, surface_flux=pair, volume_integral=th.VolumeIntegralFluxDifferencing(pair)).config(q)
    assert (c.equation, c.volume_flux, c.surface_flux) == (6, 8, 8)
    with pytest.raises(NotImplementedError):
        th.DGSEMHIP(polydeg=4, surface_flux=th.flux_chan_etal,
                    volume_integral=th.VolumeIntegralFluxDifferencing(pair)).config(q)
    with pytest.raises(NotImplementedError):  # the MHD term with quasi-1D Euler
        th.DGSEMHIP(polydeg=4, surface_flux=(th.flux_chan_etal, th.flux_nonconservative_powell),
                    volume_integral=th.VolumeIntegralFluxDifferencing(pair)).config(q)


def test_shallow_water_1d_configuration(th):
    """shallowwater_shock.jl's solver: HG with the Wintermeyer pair (DG) and
    the hydrostatic reconstruction with Audusse's term (FV, surface)."""
    eq = th.ShallowWaterEquations1D(gravity_constant=9.812, H0=1.75)
    assert eq.nvariables == 3
    hr = (th.FluxHydrostaticReconstruction(th.flux_lax_friedrichs, th.hydrostatic_reconstruction_audusse_etal),
          th.flux_nonconservative_audusse_etal)
    vol = (th.flux_wintermeyer_etal, th.flux_nonconservative_wintermeyer_etal)
    ind = th.IndicatorHennemannGassner(eq, th.LobattoLegendreBasisHIP(4), alpha_max=0.5, alpha_min=0.001,
                                       alpha_smooth=True, variable=th.waterheight_pressure)
    c = th.DGSEMHIP(polydeg=4, surface_flux=hr, volume_integral=th.VolumeIntegralShockCapturingHG(
        ind, volume_flux_dg=vol, volume_flux_fv=hr)).config(eq)
    assert (c.equation, c.volume_integral, c.volume_flux, c.fv_flux, c.surface_flux, c.indicator_variable,
            c.gravity) == (7, 2, 10, 12, 12, 0, 9.812)
    with pytest.raises(NotImplementedError):  # Fjordholm's term with the reconstruction
        th.DGSEMHIP(polydeg=4, surface_flux=(hr[0], th.flux_nonconservative_fjordholm_etal),
                    volume_integral=th.VolumeIntegralFluxDifferencing(vol)).config(eq)
    assert th.prim2cons((1.75, 1.0, 0.5), eq) == pytest.approx([1.25, 1.25, 0.5])


def test_abi_struct_layout_matches_the_header(th, tmp_path):
    """The ctypes mirrors of trixi_hip_config and trixi_hip_mesh have the C
    header's size and field offsets (a C probe compiled here with gcc)."""
    import ctypes
    import shutil
    import subprocess
    L = th._lib
    gcc = shutil.which("gcc") or shutil.which("cc")
    if gcc is None:
        pytest.skip("no C compiler")
    include = os.path.join(ROOT, "include")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "trixi_hip.h"', 'int main(void) {']
    for cname, cls in (("trixi_hip_config", L.Config), ("trixi_hip_mesh", L.MeshDesc)):
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-I", include, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, cls in (("trixi_hip_config", L.Config), ("trixi_hip_mesh", L.MeshDesc)):
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_oracle_cfg_layout_matches_its_header(oracle, tmp_path):
    """The oracle's ctypes OracleCfg has oracle_cfg's size and field offsets."""
    import ctypes
    import shutil
    import subprocess
    gcc = shutil.which("gcc") or shutil.which("cc")
    if gcc is None:
        pytest.skip("no C compiler")
    cls = oracle.OracleCfg
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "trixi_oracle.h"', 'int main(void) {',
             'printf("size %zu\\n", sizeof(oracle_cfg));']
    lines += [f'printf("{f} %zu\\n", offsetof(oracle_cfg, {f}));' for f, _ in cls._fields_]
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-I", os.path.join(ROOT, "oracle"), str(src), "-o", str(exe)], check=True)
    got = dict((l.split()[0], int(l.split()[1]))
               for l in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n") if l)
    assert got["size"] == ctypes.sizeof(cls)
    for f, _ in cls._fields_:
        assert got[f] == getattr(cls, f).offset, f


def test_reference_configurations_are_built(th):
    """Every configuration of the reference's test files (and the benchmark
    configurations) has a registered kernel in both precisions: creating the
    semidiscretization gets past the kernel lookup, which comes before any
    device allocation (without a GPU it then fails allocating, never with
    'unsupported combination')."""
    from cases import product_semi
    from reference_configs import BENCHMARK_CASES, REFERENCE_CASES
    L = th._lib
    for c in REFERENCE_CASES + BENCHMARK_CASES:
        for prec in ("fp64", "mixed"):
            try:
                product_semi(th, c, prec).close()
            except th.TrixiHipError as e:
                assert e.code != L.EUNSUPPORTED, (c.name, prec, str(e))
