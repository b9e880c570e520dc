"""MI355X-native 2-D FMM Biot-Savart velocity solver (drop-in for the FMM
velocity path of anggarapangestu/LSMPS_VPM_v2)."""
from ._lib import FP64, MIXED, VpmFmmError
from .solver import (FmmSolver, Particle, TreeCell, VelocityCalc, fmm2D, fmm3D,
                     velocity_biot_savart)

__all__ = ["FP64", "MIXED", "VpmFmmError", "FmmSolver", "Particle", "TreeCell", "fmm3D",
           "VelocityCalc", "fmm2D", "velocity_biot_savart"]
