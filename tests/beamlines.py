"""Beamlines shared by the tests."""
from trajectories.configs import reference_example


def reference_example_beamline(sign=+1.0):
    """The beamline of the reference's examples/lens_simulation_beamline.py:18-72 with the reference lens
    (integrator 'ref', linear field, a(r) table built as electrostatic_lens.py:193-209 does, here from
    the two-level Stark form)."""
    return reference_example(sign)
