"""xtrack_amd: multi-turn FP64 particle tracking on AMD Instinct MI355X.

Drop-in for the ``Line.track`` path of xsuite/xtrack: same element classes
(Drift, Multipole, Quadrupole, Bend, Sextupole, Octupole, Cavity, LimitRect,
LimitEllipse, Marker), same SoA ``Particles`` fields and the same
``track(particles, num_turns, ele_start/ele_stop, num_elements,
turn_by_turn_monitor)`` semantics, executed by a persistent HIP kernel.
"""
from .constants import PROTON_MASS_EV, ELECTRON_MASS_EV, C_LIGHT
from .particles import Particles
from .elements import (Marker, Drift, Multipole, Quadrupole, Bend, Sextupole,
                       Octupole, Cavity, LimitRect, LimitEllipse, DipoleEdge,
                       RFMultipole)
from .monitor import ParticlesMonitor
from .line import Line
from .tracker import Tracker, TrackFlags, HIPBackend
from .slicing import Strategy, Teapot, Uniform
from .batch import BatchTracker

__version__ = '0.1.0'
