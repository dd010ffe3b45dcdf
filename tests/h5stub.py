"""An in-memory stand-in for the part of h5py that trajectories.hdf and trajectories.utils use.

h5py is not installed here, so the HDF5 writer and readers run against this stand-in.  It keeps
the behaviour they rely on: ``File(path, mode)`` as a context manager over a root group that
persists per path ("r" needs an existing file); slash-separated paths with intermediate groups
created on demand; ``create_group`` raising ValueError for an existing name; ``keys()`` in name
order (h5py's default); attributes stored as numpy scalars or arrays, strings as str, and objects
rejected with TypeError as HDF5 has no type for them; datasets read back whole with ``[()]``.
What it does not check is the HDF5 file format itself.
"""
from __future__ import annotations

import types

import numpy as np


def _attr_value(v):
    if isinstance(v, str):
        return v
    a = np.asarray(v)
    if a.dtype == object:
        raise TypeError(f"Object dtype {a.dtype!r} has no native HDF5 equivalent")
    return a[()] if a.ndim == 0 else a.copy()


class _Attrs(dict):
    def __setitem__(self, key, value):
        super().__setitem__(key, _attr_value(value))


class Dataset:
    def __init__(self, data):
        self._a = np.array(data)
        self.attrs = _Attrs()

    def __getitem__(self, key):
        return self._a.copy() if key == () else self._a[key]


class Group:
    def __init__(self):
        self._children = {}
        self.attrs = _Attrs()

    @staticmethod
    def _parts(path):
        return [p for p in str(path).split("/") if p]

    def _node(self, parts, create=False):
        node = self
        for p in parts:
            if not isinstance(node, Group):
                raise KeyError("/".join(parts))
            if p not in node._children:
                if not create:
                    raise KeyError("/".join(parts))
                node._children[p] = Group()
            node = node._children[p]
        return node

    def _new(self, path, obj):
        parts = self._parts(path)
        parent = self._node(parts[:-1], create=True)
        if parts[-1] in parent._children:
            raise ValueError("Unable to create group (name already exists)")
        parent._children[parts[-1]] = obj
        return obj

    def create_group(self, path):
        return self._new(path, Group())

    def create_dataset(self, name, data=None):
        return self._new(name, Dataset(data))

    def __getitem__(self, path):
        return self._node(self._parts(path))

    def __contains__(self, path):
        try:
            self._node(self._parts(path))
        except KeyError:
            return False
        return True

    def keys(self):
        return sorted(self._children)


class _File(Group):
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def make_module():
    """A fresh module object to put in sys.modules["h5py"]; files live as long as it does."""
    files = {}

    def File(path, mode="r"):
        key = str(path)
        if mode == "r":
            if key not in files:
                raise FileNotFoundError(key)
        elif mode == "w" or key not in files:
            files[key] = _File()
        return files[key]

    mod = types.ModuleType("h5py")
    mod.File = File
    return mod
