"""GPU: trajectories.post_processing (post_processing.py:9-140 of the reference) against the oracle.

find_radial_pos_dist / find_vel_dist are evaluated on the device in one launch over all molecules of a
result, by either path:
* a run_simulation result re-walks its kept molecules with one z-grid sample at z;
* a list of Molecule objects has its stored rows resampled by ctraj_sample_rows.
Both must give the oracle's restatement of the rule (orc_zgrid_from_rows) bit for bit at z on a stored
row, between rows, inside the lens, at the lens exit, before the source and beyond the last element.
"""
import numpy as np
import pytest

from oracle import oracle as O
from trajectories import configs
from trajectories import post_processing as PP
from trajectories.compile import compile_elements
from trajectories.engine import Engine
from trajectories.molecule import Molecule, Trajectory
from trajectories.trajectory_simulator import TrajectorySimulator

pytestmark = pytest.mark.gpu

NAMES6 = ("x", "y", "vx", "vy", "vz", "t")


def oracle_at(rows, n_rows, z):
    """[6][n] of the oracle's rule at z for each molecule's rows; NaN where skipped."""
    return np.array([O.zgrid_from_rows(rows[i], int(n_rows[i]), z, 1.0, 1)[:, 0] for i in range(rows.shape[0])]).T


def z_values(bl, rows, n_rows):
    lens = bl.elements[1]
    j = int(np.nonzero(n_rows >= 8)[0][0])  # a molecule that entered the lens
    z_row = float(rows[j, 5, 2])  # its second row inside the lens
    z_mid = 0.5 * float(rows[j, 5, 2] + rows[j, 6, 2])
    return {"stored row": z_row, "between rows": z_mid, "inside lens": 1.1234567, "lens exit": lens.z1,
            "aperture 2 face": 1.6, "before the source": -0.05, "beyond the last element": 3.0}


def molecules_from_rows(rows, n_rows, labels, names):
    mols = []
    for i in range(rows.shape[0]):
        m = Molecule()
        m.trajectory = Trajectory.from_rows(rows[i, : n_rows[i]])
        m.set_aperture_hit(names[int(labels[i])])
        mols.append(m)
    return mols


def test_sample_rows_equals_the_oracle_rule():
    bl = configs.config1(stark=configs.stark_table(+1.0))
    X, V = configs.draw_source(3000, seed=91)
    eng = Engine(bl)
    rows, nrows, lab = eng.record_rows(X[0], X[1], X[2], V[0], V[1], V[2], cap=96)
    mols = molecules_from_rows(rows, nrows, lab, eng.labels)
    for name, z in z_values(bl, rows, nrows).items():
        dev, valid = PP.sample_rows(mols, z)
        ref = oracle_at(rows, nrows, z)
        np.testing.assert_array_equal(dev, ref, err_msg=name)
        np.testing.assert_array_equal(valid, ~np.isnan(ref[0]), err_msg=name)
        if name == "beyond the last element":
            assert not valid.any()
        elif name in ("stored row", "lens exit"):
            assert valid.any()


def test_find_dists_on_a_run_result_equal_the_oracle_and_the_rows_path():
    """The reference's post-processing example (lens_simulation_beamline_post_process.py:23) on an
    in-memory result: the kept molecules re-walked to z equal the oracle's rule over the oracle's
    rows, and equal the stored-rows path over the materialised molecules."""
    bl = configs.config1(stark=configs.stark_table(+1.0))
    xdist, vdist = configs.source()
    np.random.seed(92)
    sim = TrajectorySimulator()
    keep = ["Detected", "Inside lens", "Aperture 2"]
    sim.run_simulation(bl, "pp", vdist=vdist, xdist=xdist, N_traj=30_000, apertures_of_interest=keep, n_jobs=1)
    res = sim.result
    assert len(res.molecules) > 200
    X, V = configs.draw_source(30_000, seed=92)
    cb = compile_elements(bl.elements, bl.g)
    idx = res.run.aoi_index
    orc = O.propagate(cb, X[0, idx], X[1, idx], X[2, idx], V[0, idx], V[1, idx], V[2, idx], rows_cap=512)
    hits = res.molecules.aperture_hits
    assert hits == [cb.labels[k] for k in orc["label"]]
    mols = list(res.molecules)
    rows_all = orc["rows"]
    for name, z in z_values(bl, rows_all, orc["n_rows"]).items():
        for elements in (None, ["Detected"], ["Inside lens", "Aperture 2"]):
            sel = np.arange(len(hits)) if elements is None else np.array([i for i, h in enumerate(hits) if h in elements])
            ref = oracle_at(rows_all[sel], orc["n_rows"][sel], z)
            ok = ~np.isnan(ref[0])
            rho = PP.find_radial_pos_dist(res, z, elements)
            vel = PP.find_vel_dist(res, z, elements)
            tag = f"{name} elements={elements}"
            if ok.any():
                np.testing.assert_array_equal(rho, ref[0:2, ok].T, err_msg=tag)
                np.testing.assert_array_equal(vel, ref[2:5, ok].T, err_msg=tag)
            else:
                assert rho.size == 0 and vel.size == 0, tag
            # the stored-rows path over the materialised molecules gives the same bits
            vals, _ = PP.sample_at(type("R", (), {"molecules": mols})(), z, elements)
            np.testing.assert_array_equal(vals, ref[:, ok], err_msg=tag)
    # take_timestep is the reference's host helper
    x, v = PP.take_timestep(np.array([0.0, 1.0, 2.0]), np.array([1.0, 0.0, 100.0]), np.array([0.0, -9.81, 0.0]), 0.01)
    np.testing.assert_allclose(x, [0.01, 1.0 - 9.81e-4 / 2, 3.0])
    np.testing.assert_allclose(v, [1.0, -0.0981, 100.0])


def test_kept_molecules_with_lens_states_replay_their_state():
    """run_simulation with per-molecule lens states: each kept molecule's trajectory (and its sample at
    z) is the one of its own state, equal to the oracle's rows."""
    bl = configs.config4()
    xdist, vdist = configs.source()
    np.random.seed(93)
    n = 20_000
    states = np.random.RandomState(94).randint(0, 4, n).astype(np.uint8)
    np.random.seed(93)
    sim = TrajectorySimulator()
    sim.run_simulation(bl, "st", vdist=vdist, xdist=xdist, N_traj=n, apertures_of_interest=["Detected", "Inside lens"],
                       n_jobs=1, states=states)
    res = sim.result
    X, V = configs.draw_source(n, seed=93)
    eng = res.molecules._engine
    idx = res.run.aoi_index
    orc = O.propagate(eng.cb, X[0, idx], X[1, idx], X[2, idx], V[0, idx], V[1, idx], V[2, idx], state=states[idx],
                      rows_cap=512)
    assert len(res.molecules) == idx.size > 100
    assert len(np.unique(states[idx])) == 4
    for j in range(0, idx.size, max(1, idx.size // 150)):
        m = res.molecules[j]
        k = int(orc["n_rows"][j])
        tr = m.trajectory
        got = np.concatenate((tr.x, tr.v, tr.a, tr.t.reshape(-1, 1)), axis=1)
        np.testing.assert_array_equal(got, orc["rows"][j, :k], err_msg=f"kept molecule {j}")
        assert m.aperture_hit == eng.labels[int(orc["label"][j])]
    z = 1.05
    vals, sel = PP.sample_at(res, z)
    ref = oracle_at(orc["rows"], orc["n_rows"], z)
    ok = ~np.isnan(ref[0])
    np.testing.assert_array_equal(vals, ref[:, ok])
    np.testing.assert_array_equal(sel, np.nonzero(ok)[0])


def test_post_processing_of_a_result_read_back_from_hdf(monkeypatch, tmp_path):
    """The flow of the reference's lens_simulation_beamline_post_process.py:15-23: a run saved to
    HDF5 and read back with trajectories.utils gives the same distributions at z as the run in
    memory (h5py is absent here: the writer and readers run against the in-memory stand-in)."""
    import sys

    import h5stub
    from trajectories.utils import import_sim_result_from_hdf

    monkeypatch.setitem(sys.modules, "h5py", h5stub.make_module())
    bl = configs.config1(stark=configs.stark_table(+1.0))
    xdist, vdist = configs.source()
    np.random.seed(95)
    sim = TrajectorySimulator()
    sim.run_simulation(bl, "pp", vdist=vdist, xdist=xdist, N_traj=20_000, apertures_of_interest=["Detected", "Inside lens"],
                       n_jobs=1)
    res = sim.result
    mols = list(res.molecules)
    assert len(mols) > 100
    path = tmp_path / "lens_simulation_beamline.hdf"
    res.save_to_hdf(path, "run")
    back = import_sim_result_from_hdf(path, "run")
    order = sorted(range(len(mols)), key=lambda i: f"molecule_{i}")
    in_name_order = type("R", (), {"molecules": [mols[i] for i in order]})()
    assert [m.aperture_hit for m in back.molecules] == [mols[i].aperture_hit for i in order]
    for z in (back.beamline.find_element("Aperture 2").z0, 1.1234567, 2.5):
        for elements in (None, ["Detected"]):
            rho = PP.find_radial_pos_dist(back, z, elements)
            np.testing.assert_array_equal(rho, PP.find_radial_pos_dist(in_name_order, z, elements))
            np.testing.assert_array_equal(PP.find_vel_dist(back, z, elements),
                                          PP.find_vel_dist(in_name_order, z, elements))
            assert rho.shape[0] > 0 or z == 2.5
