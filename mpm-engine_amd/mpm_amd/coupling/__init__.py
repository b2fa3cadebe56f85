"""Robot-to-material coupling helpers (the reference's warpmpm.coupling): the
stress-integral reaction on a box end-effector and a force-admittance controller.
`Solver.box_contact_wrench` evaluates the same estimator on the device."""
from .admittance import ForceAdmittance
from .wrench import box_contact_wrench

__all__ = ["ForceAdmittance", "box_contact_wrench"]
