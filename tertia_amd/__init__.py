"""tertia_amd — MI355X-native quasi-static xi-slice wake sweep.

Product path: ``libqss.so`` (hand-written HIP kernels for gfx950 behind the C
ABI of ``include/qss.h``) and the host mirror :class:`tertia_amd.mesh.Mesh`
of the reference's ``Mesh`` layer-step interface.
"""
from .params import (Bunch, Driver, Params, linear_wake_64, blowout_256, two_bunch_512,  # noqa: F401
                     parameter_scan, profile_gradient_256, small, FIELDS, FIELD_KINDS,
                     QSS_KIND_C, QSS_KIND_S, QSS_SHAPE_CIC, QSS_SHAPE_NGP_REF, QSS_PREC_FP64,
                     QSS_PREC_MIXED32)
from .deck import DeckError, load_deck, load_driver  # noqa: F401


def Mesh(*args, **kwargs):
    """Create a :class:`tertia_amd.mesh.Mesh` (loads libqss.so, raises if absent)."""
    from .mesh import Mesh as _Mesh
    return _Mesh(*args, **kwargs)
