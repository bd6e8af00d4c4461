"""Host-side API: schedule parsing, lattices, model validation (CPU only)."""

import numpy as np
import pytest

from critmc_amd import Ising, chain, cubic, derive, parse_updates, square, triangular
from critmc_amd.kernels.common import UPDATE_CODES


def test_parse_updates_rows_and_errors():
    assert parse_updates("heatbath+4*overrelax+wolff") == [("heatbath", None), ("overrelax", 4), ("wolff", None)]
    assert parse_updates(" Metropolis ") == [("metropolis", None)]
    for bad in ("0*wolff", "x*metropolis", "gibbs", "", "+"):
        with pytest.raises(ValueError):
            parse_updates(bad)
    assert set(UPDATE_CODES) >= {"metropolis", "heatbath", "overrelax", "wolff", "sw"}


def test_lattice_contract():
    for lat in (chain(6), square(4, 6), cubic(4), triangular(3)):
        n, z = lat.nbr.shape
        for k in range(z):
            assert np.all(lat.nbr[lat.nbr[:, k], k ^ 1] == np.arange(n))
        assert lat.n_bonds == n * z // 2
        assert np.all(lat.bonds[:, 1] == lat.nbr[lat.bonds[:, 0], 2 * lat.bond_axis])
    lat = square(4, 6)
    assert lat.site(1, 2) == 1 + 4 * 2
    assert np.array_equal(lat.sublattice, (lat.coords.sum(axis=1) % 2).astype(np.int8))
    assert square(5).sublattice is None and triangular(4).sublattice is None


def test_model_validation():
    from critmc_amd import ON, Potts, XY

    m = Ising(square(4))
    m.check_updates(["metropolis", "metropolis-random", "heatbath", "wolff", "sw"])
    with pytest.raises(ValueError):
        m.check_updates(["overrelax"])
    # odd extents and the triangular lattice run in the reference's site orders
    for lat in (square(5), square(3), triangular(4), triangular(3)):
        for model in (Ising(lat), Potts(lat, 3)):
            model.check_updates(["metropolis", "metropolis-random", "heatbath", "wolff", "sw"])
            assert model.sweep_order([[0, 1]]) == "sequential"
        for model in (XY(lat), ON(lat, 3)):
            model.check_updates(["metropolis", "metropolis-random", "heatbath", "overrelax",
                                 "wolff", "sw"])
            assert model.sweep_order([[5, 1]]) == "sequential"
    assert Ising(square(4)).sweep_order([[0, 1], [3, 1]]) == "checkerboard"
    assert Ising(square(4)).sweep_order([[1, 1]]) == "sequential"
    assert XY(square(4)).sweep_order([[0, 1], [5, 2]]) == "checkerboard"
    assert ON(square(4), 3).sweep_order([[1, 1]]) == "sequential"
    with pytest.raises(ValueError, match="n = 2 and 3"):
        ON(square(4), 4)
    with pytest.raises(ValueError, match="zero field"):
        Ising(square(4), h=0.1).check_updates(["wolff"])
    with pytest.raises(ValueError):
        Ising(square(4), h=float("inf"))


def test_neighbour_tables_must_be_in_pair_form():
    import dataclasses

    from critmc_amd.kernels.table import check_lattice

    check_lattice(triangular(3))
    check_lattice(square(5, 3))
    lat = triangular(3)
    bad = dataclasses.replace(lat, nbr=lat.nbr[:, [2, 1, 0, 3, 4, 5]])  # +x and +y swapped
    with pytest.raises(ValueError, match="pair form"):
        check_lattice(bad)


def test_initial_state_matches_reference_generator_use():
    from critmc_amd import make_rng

    m = Ising(square(8))
    s = m.initial_state("random", make_rng(5))
    ref = (2 * make_rng(5).integers(0, 2, size=64) - 1).astype(np.int8)
    assert np.array_equal(s, ref)
    assert np.all(m.initial_state("ordered", None) == 1)
    assert m.begin(np.ones(64, np.int8)) == {"e_b": -128, "m": 64}


def test_derive_ising():
    raw = {"E": np.array([-8.0, -4.0]), "M": np.array([4, -2])}
    out = derive(raw, {"N": 4, "model": "ising"})
    assert np.allclose(out["e"], [-2.0, -1.0])
    assert np.allclose(out["m2"], [1.0, 0.25])
    assert np.allclose(out["absm"], [1.0, 0.5])
