"""HDF5 output in the reference's layout (trajectory_simulator.py:160-254, molecule.py:97-106,169-185,
apertures.py:56-80).  Requires h5py, which is optional: without it these raise ImportError.

Beyond the reference's attributes, a lens group also holds its force tables as datasets (a
``stark`` subgroup with one (E, W) knot pair per molecular state, an ``a_interp`` subgroup with the
(r, a) knots) and the beamline group holds ``g``, so that a run read back with
:mod:`trajectories.utils` propagates again.  Readers that look at attributes only, the reference's
among them, see the reference's layout."""
from __future__ import annotations

from copy import copy

import numpy as np


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
        if "stark" in skip and getattr(el, "stark", None) is not None:
            _save_splines(f, group_path + "/stark", el.stark)
        if "a_interp" in skip and getattr(el, "a_interp", None) is not None:
            t = el.a_interp
            if hasattr(t, "x") and hasattr(t, "y"):
                f.create_group(group_path + "/a_interp")
                f[group_path + "/a_interp"].create_dataset("x", data=np.asarray(t.x, dtype=np.float64))
                f[group_path + "/a_interp"].create_dataset("y", data=np.asarray(t.y, dtype=np.float64))


def _save_splines(f, path, stark):
    """W(E) knots of each state's spline: datasets E_i, W_i; attrs count and is_list."""
    splines = list(stark) if isinstance(stark, (list, tuple)) else [stark]
    f.create_group(path)
    f[path].attrs["count"] = len(splines)
    f[path].attrs["is_list"] = isinstance(stark, (list, tuple))
    for i, spl in enumerate(splines):
        f[path].create_dataset(f"E_{i}", data=np.asarray(spl.E, dtype=np.float64))
        f[path].create_dataset(f"W_{i}", data=np.asarray(spl.W, dtype=np.float64))


def load_splines(group):
    """The stark value stored by _save_splines: one WSpline or a list of them."""
    from .stark import WSpline

    out = [WSpline(group[f"E_{i}"][()], group[f"W_{i}"][()]) for i in range(int(group.attrs["count"]))]
    return out if bool(group.attrs["is_list"]) else out[0]


def save_beamline(beamline, filepath, run_name):
    h5py = _h5py()
    with h5py.File(filepath, "a") as f:
        f.create_group(run_name + "/beamline/")
        f[run_name + "/beamline"].attrs["g"] = float(beamline.g)
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
