"""MI355X-native particle tracking engine for AT lattices.

The package keeps pyat's tracking call surface (``lattice_track``,
``element_track``, ``internal_lpass``, ``MPMode``, ``gpu_info``) and runs it
on a persistent FP64 HIP kernel through the C ABI declared in
``include/atx.h``.
"""

from .lattice import (  # noqa: F401
    AtError, AtWarning, Particle, Element, LongElement, Marker, Monitor, Aperture, Drift,
    ThinMultipole, Multipole, Dipole, Bend, Quadrupole, Sextupole, Octupole, RFCavity,
    Corrector, Lattice, load_mat, e_mass, clight,
)
