"""Reading a saved run back from HDF5 (utils.py:15-159 of the reference).

The readers invert :mod:`trajectories.hdf`: a run group holds ``beamline/<element>`` groups whose
attributes are the element's fields plus its class name, ``position_distribution`` and
``velocity_distribution`` groups of the same form, a ``counter`` group whose attributes are the
counts, and ``trajectories/molecule_<i>`` groups with datasets x, v, a, t and attributes
aperture_hit, alive.  Requires h5py (optional): without it these raise ImportError.

As in the reference, groups are visited in the order h5py lists them (by name), so elements are
re-sorted by z0 by the Beamline and molecules come back in name order (molecule_0, molecule_1,
molecule_10, ...); an object is rebuilt from the attributes that name parameters of its class's
constructor.  A lens's force tables, which the reference does not store, are restored from the
datasets :mod:`trajectories.hdf` writes beside them.
"""
from __future__ import annotations

import inspect
from pathlib import Path
from typing import List

from .beamline import Beamline
from .hdf import _h5py, load_splines
from .molecule import Molecule, Trajectory
from .trajectory_simulator import Counter, SimulationResult

__all__ = ["import_element_from_hdf", "import_beamline_from_hdf", "import_trajectories_from_hdf",
           "import_distribution_from_hdf", "import_counter_from_hdf", "import_sim_result_from_hdf"]


def _init_params(cls) -> List[str]:
    return list(inspect.signature(cls.__init__).parameters)[1:]


def import_element_from_hdf(element_name: str, filepath: Path, run_name: str):
    """The beamline element saved under run_name/beamline/element_name (utils.py:15-42)."""
    from . import beamline_elements

    h5py = _h5py()
    path = run_name + "/beamline/" + element_name
    with h5py.File(filepath, "r") as f:
        attributes = dict(f[path].attrs.items())
        group = f[path]
        stark = load_splines(group["stark"]) if "stark" in group else None
        table = (group["a_interp"]["x"][()], group["a_interp"]["y"][()]) if "a_interp" in group else None
    cls = getattr(beamline_elements, attributes.pop("class"))
    names = set(_init_params(cls)) & set(attributes)
    kwargs = {key: attributes[key] for key in names}
    if stark is not None:
        kwargs["stark"] = stark
    if table is not None:
        from .beamline_elements.electrostatic_lens import LinearTable

        kwargs["a_interp"] = LinearTable(*table)
    return cls(**kwargs)


def import_beamline_from_hdf(filepath: Path, run_name: str) -> Beamline:
    """The run's beamline, one element per group of run_name/beamline (utils.py:45-62)."""
    h5py = _h5py()
    with h5py.File(filepath, "r") as f:
        group = f[run_name + "/beamline"]
        names = list(group.keys())
        g = group.attrs["g"] if "g" in group.attrs else None
    elements = [import_element_from_hdf(name, filepath, run_name) for name in names]
    return Beamline(elements) if g is None else Beamline(elements, float(g))


def import_trajectories_from_hdf(filepath: Path, run_name: str, group_name: str = "trajectories") -> List[Molecule]:
    """One Molecule per group of run_name/group_name, with its stored rows, label and alive flag
    (utils.py:65-109)."""
    h5py = _h5py()
    molecules = []
    with h5py.File(filepath, "r") as f:
        parent = f[run_name + "/" + group_name]
        for name in list(parent.keys()):
            group = parent[name]
            tr = Trajectory(n_rows=0)
            tr.x = group["x"][()]
            tr.v = group["v"][()]
            tr.a = group["a"][()]
            tr.t = group["t"][()]
            tr.n = int(tr.t.shape[0])
            mol = Molecule()
            mol.trajectory = tr
            mol.set_aperture_hit(group.attrs["aperture_hit"])
            mol.alive = group.attrs["alive"]
            molecules.append(mol)
    return molecules


def import_distribution_from_hdf(dist_name: str, filepath: Path, run_name: str):
    """The distribution saved under run_name/dist_name; every constructor parameter must be stored
    (utils.py:112-136)."""
    from . import distributions

    h5py = _h5py()
    with h5py.File(filepath, "r") as f:
        attributes = dict(f[run_name + "/" + dist_name].attrs.items())
    cls = getattr(distributions, attributes.pop("class"))
    return cls(**{key: attributes[key] for key in _init_params(cls)})


def import_counter_from_hdf(filepath: Path, run_name: str) -> Counter:
    """A Counter whose counts are the attributes of run_name/counter (utils.py:139-150)."""
    h5py = _h5py()
    with h5py.File(filepath, "r") as f:
        attributes = dict(f[run_name + "/counter"].attrs.items())
    counter = Counter()
    counter.counter_dict = attributes
    return counter


def import_sim_result_from_hdf(filepath: Path, run_name: str) -> SimulationResult:
    """The SimulationResult saved as run_name (utils.py:153-159)."""
    beamline = import_beamline_from_hdf(filepath, run_name)
    xdist = import_distribution_from_hdf("position_distribution", filepath, run_name)
    vdist = import_distribution_from_hdf("velocity_distribution", filepath, run_name)
    counter = import_counter_from_hdf(filepath, run_name)
    molecules = import_trajectories_from_hdf(filepath, run_name)
    return SimulationResult(counter, beamline, xdist, vdist, molecules)
