"""pe_amd: MI355X-native HCSITS rigid-body time step (HIP, gfx950, FP64).

The hot path of pe's rigid-body simulation -- hash-grid broadphase, sphere/box/plane
contacts, NBGS relaxation with InelasticCoulombContactByDecoupling and semi-implicit
integration -- runs as hand-written HIP kernels behind the C-ABI in include/pe_hip.h.
``World`` mirrors pe's World/CollisionSystem API on top of it.
"""
from . import _abi
from ._abi import PeHipError, load
from .world import (World, InelasticFrictionlessContact, ApproximateInelasticCoulombContactByDecoupling,
                    InelasticCoulombContactByDecoupling)

SWEEP_COLOURED = _abi.SWEEP_COLOURED
SWEEP_REFERENCE_ORDER = _abi.SWEEP_REFERENCE_ORDER


def version():
    return load().pehip_version().decode()


__all__ = ["World", "PeHipError", "load", "version", "SWEEP_COLOURED", "SWEEP_REFERENCE_ORDER",
           "InelasticFrictionlessContact", "ApproximateInelasticCoulombContactByDecoupling",
           "InelasticCoulombContactByDecoupling"]
