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

from .tracking import (  # noqa: E402,F401
    End, MPMode, atpass, diffusion_matrix, element_pass, element_track, gpu_core_count, gpu_info,
    internal_epass, internal_lpass, internal_plpass, lattice_pass, lattice_track, patpass,
)
from . import acceptance as _acceptance  # noqa: E402
from . import physics as _physics  # noqa: E402
from .acceptance import (  # noqa: E402,F401
    GridMode, get_1d_acceptance, get_acceptance, get_horizontal_acceptance,
    get_momentum_acceptance, get_vertical_acceptance,
)
from .physics import (  # noqa: E402,F401
    DConstant, ELossMethod, element_matrices, find_elem_m66, find_m44, find_m66, find_orbit,
    find_orbit4, find_orbit4_batch, find_orbit6, find_orbit6_batch, find_sync_orbit,
    get_energy_loss, get_timelag_fromU0,
)

# Lattice methods, as pyat attaches them (physics/__init__.py, acceptance)
for _name in ("find_orbit", "find_orbit4", "find_orbit6", "find_sync_orbit", "find_m44",
              "find_m66", "get_energy_loss", "get_timelag_fromU0"):
    setattr(Lattice, _name, getattr(_physics, _name))
for _name in ("get_acceptance", "get_1d_acceptance", "get_horizontal_acceptance",
              "get_vertical_acceptance", "get_momentum_acceptance"):
    setattr(Lattice, _name, getattr(_acceptance, _name))
from .formats import load_json, load_lattice, load_repr, save_json, save_lattice, save_repr  # noqa: E402,F401
