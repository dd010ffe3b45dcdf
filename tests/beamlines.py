"""Beamlines shared by the tests."""
from trajectories.beamline import Beamline
from trajectories.beamline_elements import CircularAperture, ElectrostaticLens, FieldPlates, RectangularAperture
from trajectories.stark import TwoLevelStark


def reference_example_beamline(sign=+1.0):
    """The beamline of the reference's examples/lens_simulation_beamline.py:18-72 with the reference lens
    (integrator 'ref', linear field, a(r) table built as electrostatic_lens.py:193-209 does, here from
    the two-level Stark form)."""
    m = 0.0254
    fourK = CircularAperture(z0=1.7 * m, L=0.25 * m, d=1 * m, name="4K shield")
    fortyK = CircularAperture(z0=fourK.z1 + 1.25 * m, L=0.25 * m, d=1 * m, name="40K shield")
    bb_exit = CircularAperture(z0=fortyK.z1 + 2.5 * m, L=0.75 * m, d=4 * m, name="BB exit")
    lens = ElectrostaticLens(z0=bb_exit.z1 + 33 * m, L=0.6, name="ES lens", state=TwoLevelStark(sign=sign))
    plates = FieldPlates(z0=2.43, L=3.0, w=0.02, name="Field plates")
    dr = RectangularAperture(z0=plates.z1 + 39.9 * m, L=0.25 * m, name="DR aperture", w=0.018, h=0.03)
    return Beamline([fourK, fortyK, bb_exit, lens, plates, dr])
