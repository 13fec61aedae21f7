"""fastmultipole_amd: MI355X-native FMM evaluator behind FastMultipole's fmm!/direct! interface.

The GPU library (libfmmhip.so, C ABI in include/fmmhip.h) does the tree build,
interaction lists, P2P, P2M/M2M/M2L/L2L/L2P; this package is the host-side
mirror of the reference's entry points and user-system interface.
"""
from . import _lib
from .core import (Context, Evaluation, DerivativesSwitch, ErrorMethod, InteractionListMethod, PowerAbsoluteGradient,
                   PowerAbsolutePotential, PowerRelativeGradient, PowerRelativePotential,
                   RotatedCoefficientsAbsoluteGradient, RotatedCoefficientsRelativeGradient, Barba, SelfTuning,
                   SelfTuningTreeStop, default_context, direct, fmm, fmm_batch, make_options, FastGaussSeidel, Server,
                   solve, tune_fmm, autotune)
from .systems import (DipolePanels, DipolePoints, GaussianVortexParticles, Gravitational, SourceDipolePanels, SourcePanels,
                      SourceVortexPoints,
                      VortexParticles)

__all__ = ["Context", "Evaluation", "DerivativesSwitch", "ErrorMethod", "InteractionListMethod",
           "PowerAbsoluteGradient", "PowerAbsolutePotential", "PowerRelativeGradient", "PowerRelativePotential",
           "RotatedCoefficientsAbsoluteGradient", "RotatedCoefficientsRelativeGradient", "Barba", "SelfTuning",
           "SelfTuningTreeStop", "default_context", "direct", "fmm", "fmm_batch", "make_options", "Gravitational",
           "VortexParticles", "DipolePoints", "GaussianVortexParticles", "SourceVortexPoints", "SourcePanels", "DipolePanels", "SourceDipolePanels",
           "FastGaussSeidel", "Server", "solve", "tune_fmm", "autotune"]
