"""Design-side arithmetic (no GPU): the oracle's literal restatements pinned by
closed forms, and the product's vectorised / torch versions against them.

Pins of the oracle: PtoFderivative equals a central difference of
SmoothDensity; hand-built weights r - distance; schedule end points; the
first Adam step of a constant gradient; exact interpolation of a linear
table.  A restatement that dropped a term, swapped the two branches of the
projection or mirrored the wrong axis fails one of them.
"""
import math

import numpy as np
import pytest
import torch

import design_oracle as D
import dda_oracle as O
from dda_model import density, design, materials


# ---------------------------------------------------------------------------
# oracle pins
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("x", [0.05, 0.3, 0.5, 0.62, 0.97])
@pytest.mark.parametrize("beta,ita", [(1.0, 0.5), (8.0, 0.4), (30.0, 0.6)])
def test_smooth_density_derivative_is_central_difference(x, beta, ita):
    h = 1e-6
    fd = (D.smooth_density(x + h, beta, ita) - D.smooth_density(x - h, beta, ita)) / (2 * h)
    assert D.ptof_derivative(x, beta, ita) == pytest.approx(fd, rel=1e-6, abs=1e-8)
    # end points map to themselves and ita is a fixed point
    assert D.smooth_density(0.0, beta, ita) == pytest.approx(0.0, abs=1e-15)
    assert D.smooth_density(1.0, beta, ita) == pytest.approx(1.0, abs=1e-15)
    assert D.smooth_density(ita, beta, ita) == pytest.approx(ita, abs=1e-15)


def test_free_weights_hand_built():
    # 3 x 1 x 1 row of dipoles, one parameter each, radius 1.5: each dipole
    # sees itself (w = 1.5) and its neighbours at distance 1 (w = 0.5)
    geo = [0, 0, 0, 1, 0, 0, 2, 0, 0]
    fw = D.free_weights(geo, [[0], [1], [2]], 1.5, False, 0, 0)
    assert sorted(fw[1]) == sorted([(0.5, 0), (1.5, 1), (0.5, 2)])
    assert sorted(fw[0]) == sorted([(1.5, 0), (0.5, 1)])
    # the circle test is r^2 - 0.01: distance exactly r is outside
    assert D.free_weights(geo, [[0], [1], [2]], 1.0, False, 0, 0)[1] == [(1.0, 1)]
    # periodic images at +-Lm (Lm = 3 cells) make the row a ring (Ln far away)
    fwp = D.free_weights(geo, [[0], [1], [2]], 1.5, True, 3, 100)
    assert sorted(fwp[0]) == sorted([(1.5, 0), (0.5, 1), (0.5, 2)])


def test_schedule_end_points():
    assert D.exp_update(0, 200, 1, 50) == pytest.approx(1)
    assert D.exp_update(200, 200, 1, 50) == pytest.approx(50)
    assert D.exp_update(301, 400, 1, 50) == 0.5            # the reference's constant
    assert D.piecewise_update(100, 200, 0, 50) == 0
    assert D.piecewise_update(130, 200, 0, 50) == pytest.approx(10)
    assert D.piecewise_update(150, 200, 0, 50) == pytest.approx(20)
    assert D.piecewise_update(161, 200, 0, 50) == 50
    assert D.linear_update(50, 200, 0, 40) == pytest.approx(10)


def test_adam_first_step_and_table_interpolation():
    g = [0.3, -2.0, 1e-3]
    step = D.Adam()(g)
    # t = 1: V = g, S = g^2 exactly (the bias corrections cancel), step = g / (|g| + 1e-8)
    assert step == pytest.approx([x / (abs(x) + 1e-8) for x in g], rel=1e-12)
    rows = [(400e-9 + 50e-9 * k, 2.0 + 1e6 * (400e-9 + 50e-9 * k)) for k in range(9)]
    assert D.get_material_from_rows(rows, rows, 533e-9).real == pytest.approx(2.0 + 0.533, rel=1e-12)
    with pytest.raises(ValueError):
        D.get_material_from_rows(rows, rows, 900e-9)


# ---------------------------------------------------------------------------
# product against the oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("symmetry,periodic", [("None", False), ("4fold", False), ("4fold", True)])
def test_filter_weights_and_map_match_oracle(symmetry, periodic):
    nx, ny, nz = 8, 8, 3
    geo = O.generate_geometry(nx, ny, nz)
    nf_o, gpara_o, p2g = O.parameter_mapping(geo, nx * ny * nz, nz, symmetry, [3.5, 3.5])
    nf, gpara = density.parameter_map(geo, nz, symmetry, (3.5, 3.5))
    assert nf == nf_o and np.array_equal(gpara, gpara_o)
    W = density.filter_weights(geo, gpara, nf, 2.3, periodic, nx, ny)
    fw = D.free_weights(list(geo), p2g, 2.3, periodic, nx, ny)
    Wo = np.zeros((nf, nf))
    for i, row in enumerate(fw):
        for w, j in row:
            Wo[i, j] += w
    assert np.allclose(W, Wo, rtol=1e-14, atol=1e-14)
    rng = np.random.default_rng(1)
    origin = rng.uniform(0, 1, nf)
    filt = density.cone_filter_params(torch.tensor(origin), W).numpy()
    assert np.allclose(filt, D.filtered_parameters(fw, origin), rtol=1e-13)
    beta, ita = 6.0, 0.45
    g = rng.standard_normal(nf)
    got = density.gradients_filtered_reference(g, filt, W, beta, ita)
    assert np.allclose(got, D.gradients_filtered(list(g), list(filt), fw, beta, ita), rtol=1e-12)


def test_projection_and_schedules_match_oracle():
    x = torch.linspace(0, 1, 101, dtype=torch.float64)
    for beta, ita in [(1.0, 0.5), (12.0, 0.3)]:
        got = density.smooth_density(x, beta, ita).numpy()
        assert np.allclose(got, [D.smooth_density(float(v), beta, ita) for v in x], rtol=1e-13, atol=1e-15)
        d = density.smooth_density_derivative(x, beta, ita).numpy()
        assert np.allclose(d, [D.ptof_derivative(float(v), beta, ita) for v in x], rtol=1e-13)
    with pytest.raises(ValueError):
        density.smooth_density(torch.tensor([1.2], dtype=torch.float64), 1.0, 0.5)
    for it in (0, 37, 120, 250, 301):
        assert density.beta_at("exp", it, 300, 1, 40) == pytest.approx(D.exp_update(it, 300, 1, 40))
        assert density.beta_at("piecewise", it, 300, 1, 40) == pytest.approx(D.piecewise_update(it, 300, 1, 40))
        assert density.beta_at("linear", it, 300, 1, 40) == pytest.approx(D.linear_update(it, 300, 1, 40))
    with pytest.raises(RuntimeError):
        density.beta_at("cosine", 1, 10, 0, 1)


def test_autograd_of_projection_is_ptof():
    x = torch.tensor([0.1, 0.45, 0.8], dtype=torch.float64, requires_grad=True)
    density.smooth_density(x, 7.0, 0.45).sum().backward()
    assert np.allclose(x.grad.numpy(), [D.ptof_derivative(v, 7.0, 0.45) for v in (0.1, 0.45, 0.8)], rtol=1e-12)


def test_optimizers_and_penalties_match_oracle():
    rng = np.random.default_rng(2)
    a, o = design.AdamOptimizer(), D.Adam()
    for _ in range(5):
        g = rng.standard_normal(7)
        assert np.allclose(a(g), o(list(g)), rtol=1e-12)
    p = rng.uniform(0, 1, 50)
    cfgs = {"gaussian": {"sigma": 0.2, "mu": 0.5}, "triangular": {"slope": 1.0}, "parabolic": {}}
    for kind, cfg in cfgs.items():
        f, gf = design.PENALTIES[kind]
        assert np.allclose(f(p, cfg), D.penalty(kind, p, cfg), rtol=1e-13)
        assert np.allclose(gf(p, cfg), D.penalty_gradient(kind, p, cfg), rtol=1e-13)
    ag = design.AdaGradOptimizer()
    g1, g2 = rng.standard_normal(3), rng.standard_normal(3)
    ag(g1)
    assert np.allclose(ag(g2), g2 / (g1 * g1 + g2 * g2 + 1), rtol=1e-14)
    c = {"iter_start": 0, "iter_end": 400, "coeff_min": 0.0, "coeff_max": 1.0, "base": 100}
    assert design.coeff_linear(100, c) == pytest.approx(0.25)
    assert design.coeff_exp(400, c) == pytest.approx(1.0) and design.coeff_exp(0, c) == pytest.approx(0.0)


def test_geometry_file_round_trip(tmp_path, golden_inputs):
    _, geo = golden_inputs
    (nx, ny, nz), g = design.read_geometry(__import__("conftest").GOLDEN + "/geometry.txt")
    assert (nx, ny, nz) == (22, 22, 10) and np.array_equal(g, geo[4:])
    design.write_geometry(tmp_path / "g.txt", nx, ny, nz, g)
    assert (tmp_path / "g.txt").read_text() == open(__import__("conftest").GOLDEN + "/geometry.txt").read()
    d = design.read_dielectrics(__import__("conftest").GOLDEN + "/dielectrics.txt")
    design.write_dielectrics(tmp_path / "d.txt", d)
    assert np.array_equal(design.read_dielectrics(tmp_path / "d.txt"), d)


def test_materials():
    # fused silica n_D = 1.4585 (589.3 nm), sapphire n_o = 1.7682
    assert math.sqrt(materials.permittivity("SiO2", 589.3).real) == pytest.approx(1.45846, abs=2e-4)
    assert math.sqrt(materials.permittivity("Al2O3", 0.5893, "um").real) == pytest.approx(1.76817, abs=2e-4)
    assert materials.permittivity("Air", 500) == 1.0
    assert materials.permittivity(5.96 + 3.89j, 500) == 5.96 + 3.89j
    with pytest.raises(ValueError):
        materials.permittivity("Unobtainium", 500)


def test_table_material_matches_reference_scan(tmp_path):
    rng = np.random.default_rng(3)
    wl = np.sort(rng.uniform(300e-9, 900e-9, 40))
    re = np.column_stack((wl, rng.uniform(-10, 5, 40)))
    im = np.column_stack((wl, rng.uniform(0, 3, 40)))
    np.savetxt(tmp_path / "XRe_eps.txt", re)
    np.savetxt(tmp_path / "XIm_eps.txt", im)
    materials.register_table("X", str(tmp_path / "X"))
    for w in rng.uniform(wl[0] + 1e-12, wl[-1] - 1e-12, 25):
        assert materials.permittivity("X", w, "m") == pytest.approx(
            D.get_material_from_rows(re.tolist(), im.tolist(), w), rel=1e-12)
    with pytest.raises(ValueError):
        materials.permittivity("X", 1000.0, "nm")
