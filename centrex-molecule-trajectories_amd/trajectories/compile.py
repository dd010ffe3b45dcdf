"""Compile a Beamline into the flat, z0-sorted element table of the C ABI.

Labels are strings as in the reference's Counter: element names for apertures and field plates,
"Lens entrance" / "Inside lens" for lenses (electrostatic_lens.py:63,117) and "Detected" for
survivors (beamline.py:34-35).  Equal strings share one label id.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

LABEL_LENS_ENTRANCE = "Lens entrance"
LABEL_INSIDE_LENS = "Inside lens"
LABEL_DETECTED = "Detected"


class Registry:
    """Collects labels and lens tables while elements pack themselves."""

    def __init__(self):
        self.labels: list[str] = []
        self._ids: dict[str, int] = {}
        self.rtables: list[tuple[np.ndarray, np.ndarray]] = []
        self._rt_keys: dict[bytes, int] = {}
        self.wsplines = None  # list of WSpline (one per state)

    def label(self, name: str) -> int:
        name = str(name)
        if name not in self._ids:
            if len(self.labels) >= N.MAX_LABELS:
                raise ValueError(f"more than {N.MAX_LABELS} distinct labels")
            self._ids[name] = len(self.labels)
            self.labels.append(name)
        return self._ids[name]

    def rtable(self, r, a) -> int:
        r = np.ascontiguousarray(r, dtype=np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64)
        key = hashlib.sha1(r.tobytes() + b"|" + a.tobytes()).digest()
        if key not in self._rt_keys:
            if len(self.rtables) >= 8:
                raise ValueError("at most 8 distinct lens a(r) tables per beamline")
            self._rt_keys[key] = len(self.rtables)
            self.rtables.append((r, a))
        return self._rt_keys[key]

    def wspline(self, splines) -> None:
        if not isinstance(splines, (list, tuple)):
            splines = [splines]
        splines = list(splines)
        if self.wsplines is None:
            nseg, dE = splines[0].n_seg, splines[0].dE
            for s in splines:
                if s.n_seg != nseg or s.dE != dE:
                    raise ValueError("all state tables must share the |E| grid")
            if len(splines) > N.MAX_STATES:
                raise ValueError(f"at most {N.MAX_STATES} state tables")
            self.wsplines = splines
        elif [id(s) for s in self.wsplines] != [id(s) for s in splines]:
            raise ValueError("all spline-force lenses of a beamline must use the same state tables")


@dataclass
class CompiledBeamline:
    elems: object  # ctypes array of N.Elem
    n_elems: int
    labels: list
    phys: N.Phys
    rtables: list
    wsplines: list | None
    force: int
    integrator: int
    elements: list = field(default_factory=list)
    key: bytes = b""

    @property
    def n_labels(self):
        return len(self.labels)

    @property
    def label_detected(self):
        return self.phys.label_detected

    @property
    def n_states(self):
        return self.phys.n_states

    def label_id(self, name: str) -> int:
        return self.labels.index(name)

    def wspline_coef(self) -> np.ndarray | None:
        if not self.wsplines:
            return None
        return np.ascontiguousarray(np.stack([s.coef for s in self.wsplines]), dtype=np.float64)


def compile_elements(elements, g: float, n_states: int | None = None, max_lens_steps: int = 0) -> CompiledBeamline:
    elements = list(elements)
    if len(elements) > N.MAX_ELEMS:
        raise ValueError(f"at most {N.MAX_ELEMS} elements")
    for a, b in zip(elements, elements[1:]):
        if b.z0 < a.z0:
            raise ValueError("elements must be sorted by z0 (Beamline sorts them)")
    reg = Registry()
    arr = (N.Elem * max(1, len(elements)))()
    force = integ = None
    for i, el in enumerate(elements):
        pack = getattr(el, "_ctraj_pack", None)
        if pack is None:
            raise TypeError(f"{type(el).__name__} has no device implementation")
        E = pack(reg)
        arr[i] = E
        if E.kind == N.LENS:
            if force is not None and (force, integ) != (E.force, E.integrator):
                raise ValueError("all lenses of a beamline must share force model and integrator")
            force, integ = E.force, E.integrator
    det = reg.label(LABEL_DETECTED)
    ns = len(reg.wsplines) if reg.wsplines else 1
    if n_states is not None:
        if reg.wsplines and n_states > len(reg.wsplines):
            raise ValueError("n_states exceeds the lens state tables")
        ns = int(n_states)
    phys = N.Phys(g=float(g), n_labels=len(reg.labels), label_detected=det, n_states=ns,
                  max_lens_steps=int(max_lens_steps))
    h = hashlib.sha1(bytes(arr) + bytes(phys))
    for r, a in reg.rtables:
        h.update(r.tobytes())
        h.update(a.tobytes())
    if reg.wsplines:
        for s in reg.wsplines:
            h.update(s.coef.tobytes())
    return CompiledBeamline(arr, len(elements), reg.labels, phys, reg.rtables, reg.wsplines,
                            N.FORCE_WSPLINE_HEX if force is None else force,
                            N.INTEG_RK4 if integ is None else integ, elements, h.digest())
