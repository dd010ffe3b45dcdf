"""HDF5 output in the reference's layout (trajectory_simulator.py:160-254, molecule.py:97-106,169-185,
apertures.py:56-80).  Requires h5py, which is optional: without it these raise ImportError."""
from __future__ import annotations

from copy import copy


def _h5py():
    try:
        import h5py  # noqa: F401
    except ImportError as e:  # pragma: no cover - depends on the environment
        raise ImportError("HDF5 output needs h5py, which is not installed") from e
    import h5py

    return h5py


def save_element(el, filepath, parent_group_path, skip=(), repr_keys=()):
    h5py = _h5py()
    with h5py.File(filepath, "a") as f:
        group_path = parent_group_path + "/" + el.name
        try:
            f.create_group(group_path)
        except ValueError:
            print("Can't save beamline element. Group already exists!")
            return
        f[group_path].attrs["class"] = type(el).__name__
        for key, value in vars(el).items():
            if key in skip or value is None:
                continue
            f[group_path].attrs[key] = repr(value) if key in repr_keys else value


def save_beamline(beamline, filepath, run_name):
    h5py = _h5py()
    with h5py.File(filepath, "a") as f:
        f.create_group(run_name + "/beamline/")
    for element in beamline.elements:
        element.save_to_hdf(filepath, run_name + "/beamline/")


def save_distribution(dist, filepath, run_name):
    h5py = _h5py()
    from .distributions import CeNTREXVelocityDistribution

    group = "/velocity_distribution" if isinstance(dist, CeNTREXVelocityDistribution) else "/position_distribution"
    with h5py.File(filepath, "a") as f:
        f.create_group(run_name + group)
        f[run_name + group].attrs["class"] = type(dist).__name__
        for key, value in vars(dist).items():
            f[run_name + group].attrs[key] = value


def save_counter(counter, filepath, run_name):
    h5py = _h5py()
    with h5py.File(filepath, "a") as f:
        try:
            f.create_group(run_name + "/counter")
        except ValueError:
            raise ValueError("Can't save counter. Group already exists!")
        for key, value in counter.counter_dict.items():
            f[run_name + "/counter"].attrs[key] = value


def save_result(result, filepath, run_name):
    h5py = _h5py()
    with h5py.File(filepath, "a") as f:
        if run_name in f:
            raise ValueError(f"run {run_name!r} already exists in {filepath}")
        f.create_group(run_name)
    attrs = copy(vars(result))
    attrs.pop("molecules")
    attrs.pop("run", None)
    for value in attrs.values():
        value.save_to_hdf(filepath, run_name)
    with h5py.File(filepath, "a") as f:
        for i, molecule in enumerate(result.molecules):
            molecule.save_to_hdf(f, run_name, f"trajectories/molecule_{i}")
