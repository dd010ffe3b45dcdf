"""HDF5 save and load of a SimulationResult (trajectory_simulator.py:160-254, utils.py:15-159 of the
reference), CPU only.  h5py is not installed, so the writer and readers run against the in-memory
stand-in of tests/h5stub.py: this checks what is written where and that reading it back rebuilds
the same result; the HDF5 file format itself is not exercised (parity with h5py files unpinned)."""
import dataclasses
import sys

import numpy as np
import pytest

import h5stub
from trajectories import configs
from trajectories.beamline import Beamline
from trajectories.beamline_elements import (CircularAperture, ElectrostaticLens, FieldPlates,
                                            RectangularAperture)
from trajectories.compile import compile_elements
from trajectories.distributions import CeNTREXPositionDistribution, CeNTREXVelocityDistribution
from trajectories.molecule import Molecule
from trajectories.trajectory_simulator import Counter, SimulationResult


@pytest.fixture
def h5(monkeypatch):
    mod = h5stub.make_module()
    monkeypatch.setitem(sys.modules, "h5py", mod)
    return mod


def _result(beamline, n_mol=12):
    counter = Counter()
    counter.counter_dict = {"Aperture 1": 5, "Inside lens": 4, "Detected": 3}
    rs = np.random.RandomState(7)
    molecules = []
    for i in range(n_mol):
        m = Molecule()
        m.init_trajectory(beamline, x0=rs.normal(0, 1e-3, 3), v0=np.array([1.0, -2.0, 180.0]) + rs.normal(0, 1, 3))
        for _ in range(3 + i % 4):
            m.update_trajectory(1e-3 * (1 + rs.rand()))
        m.set_aperture_hit(["Aperture 1", "Inside lens", "Detected"][i % 3])
        if i % 3 != 2:
            m.set_dead()
        molecules.append(m)
    return SimulationResult(counter, beamline, CeNTREXPositionDistribution(d=0.015, z=0.01),
                            CeNTREXVelocityDistribution(vz=190.0, sigmaz=15.0), molecules)


def _config_beamline():
    lens = ElectrostaticLens("ES lens", 0.8, 0.6, d=0.035, V=30e3, stark=configs.config4_tables(), field="hexapole",
                             integrator="rk4", dt=1e-5)
    return Beamline([
        CircularAperture("Aperture 1", 0.3, 0.0, d=0.012),
        lens,
        FieldPlates("Plates", 1.5, 0.3, x0=0.0005, w=0.02),
        RectangularAperture("Rect", 2.0, 0.1, x0=0.001, y0=-0.002, w=0.02, h=0.03),
        CircularAperture("Detector", 2.5, 0.0, d=0.008),
    ], 9.7)


def _fields(el):
    return {f.name: getattr(el, f.name) for f in dataclasses.fields(el) if f.name not in ("stark", "a_interp", "state")}


@pytest.mark.parametrize("make", [_config_beamline, configs.reference_example], ids=["config", "reference_example"])
def test_result_round_trip(h5, make, tmp_path):
    from trajectories.utils import import_sim_result_from_hdf

    result = _result(make())
    path = tmp_path / "run.hdf"
    result.save_to_hdf(path, "run 1")
    back = import_sim_result_from_hdf(path, "run 1")

    assert {k: int(v) for k, v in back.counter.counter_dict.items()} == result.counter.counter_dict
    assert back.xdist == result.xdist and back.vdist == result.vdist
    assert back.beamline.g == result.beamline.g
    assert [type(e) for e in back.beamline.elements] == [type(e) for e in result.beamline.elements]
    for a, b in zip(back.beamline.elements, result.beamline.elements):
        assert _fields(a) == _fields(b)
    # the device sees the same element table, labels and lens tables
    g = result.beamline.g
    assert compile_elements(back.beamline.elements, g).key == compile_elements(result.beamline.elements, g).key

    # molecules come back in name order, as h5py lists groups: molecule_0, molecule_1, molecule_10, ...
    order = sorted(range(len(result.molecules)), key=lambda i: f"molecule_{i}")
    assert order[:4] == [0, 1, 10, 11]
    assert len(back.molecules) == len(result.molecules)
    for j, i in enumerate(order):
        m, o = back.molecules[j], result.molecules[i]
        for key in ("x", "v", "a", "t"):
            np.testing.assert_array_equal(getattr(m.trajectory, key), getattr(o.trajectory, key))
        assert m.aperture_hit == o.aperture_hit and bool(m.alive) == o.alive
        np.testing.assert_array_equal(m.x(), o.x())
        assert m.t() == o.t()


def test_lens_tables_round_trip(h5, tmp_path):
    from trajectories.utils import import_element_from_hdf

    bl = _config_beamline()
    path = tmp_path / "bl.hdf"
    bl.save_to_hdf(path, "r")
    f = h5.File(path, "r")
    assert set(f["r/beamline"].keys()) == {e.name for e in bl.elements}
    assert f["r/beamline/ES lens"].attrs["class"] == "ElectrostaticLens"
    assert "stark" not in f["r/beamline/ES lens"].attrs  # the reference's attributes only
    lens = import_element_from_hdf("ES lens", path, "r")
    orig = bl.elements[1]
    assert isinstance(lens.stark, list) and len(lens.stark) == len(orig.stark)
    for a, b in zip(lens.stark, orig.stark):
        np.testing.assert_array_equal(a.coef, b.coef)
        assert a.dE == b.dE

    ref = configs.reference_example()
    ref.save_to_hdf(path, "ref")
    lens = import_element_from_hdf("ES lens", path, "ref")
    r0, a0 = ref.find_element("ES lens").radial_table()
    r1, a1 = lens.radial_table()
    np.testing.assert_array_equal(r0, r1)
    np.testing.assert_array_equal(a0, a1)
    assert isinstance(lens.state, str)  # stored as its repr, as the reference does


def test_writer_errors(h5, tmp_path):
    result = _result(_config_beamline(), n_mol=2)
    path = tmp_path / "e.hdf"
    result.save_to_hdf(path, "run")
    with pytest.raises(ValueError):
        result.save_to_hdf(path, "run")
    from trajectories.utils import import_sim_result_from_hdf

    with pytest.raises(FileNotFoundError):
        import_sim_result_from_hdf(tmp_path / "missing.hdf", "run")


def test_without_h5py(monkeypatch, tmp_path):
    monkeypatch.setitem(sys.modules, "h5py", None)
    from trajectories.utils import import_counter_from_hdf

    with pytest.raises(ImportError):
        import_counter_from_hdf(tmp_path / "x.hdf", "run")
    with pytest.raises(ImportError):
        _result(_config_beamline(), n_mol=1).save_to_hdf(tmp_path / "x.hdf", "run")
