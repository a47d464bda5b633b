"""MI355X (gfx950) DQMC local sweep behind MonteCarlo.jl's DQMC / HubbardModel / run! API.

The compute path is libdqmc_hip.so (C ABI in include/dqmc_hip.h); this package holds the
host-side mirror of the reference interface: lattices, the Hubbard model and its hopping
matrix, DQMC parameters, Hirsch fields, the run! driver and equal-time measurements.
"""
from .lattice import Chain, SquareLattice, CubicLattice, Honeycomb, TriangularLattice, Lattice
from .model import (HubbardModel, AttractiveHubbardModel, RepulsiveHubbardModel,
                    HubbardModelAttractive, HubbardModelRepulsive)
from .parameters import DQMCParameters, generate_chunks
from .fields import DensityHirschField, MagneticHirschField, hirsch_alpha
from .dqmc import DQMC, run, ExitCode, MagnitudeStats, DQMCAnalysis
from .updates import (LocalSweep, GlobalFlip, GlobalShuffle, SpatialShuffle, TemporalShuffle, Denoise,
                      DenoiseFlip, StaggeredDenoise, SimpleScheduler, AcceptanceStatistics)
from .fileio import ConfigRecorder, Discarder, save, load, resume
from . import measurements

__all__ = ["Chain", "SquareLattice", "CubicLattice", "Honeycomb", "TriangularLattice", "Lattice",
           "HubbardModel", "AttractiveHubbardModel", "RepulsiveHubbardModel",
           "HubbardModelAttractive", "HubbardModelRepulsive", "DQMCParameters", "generate_chunks",
           "DensityHirschField", "MagneticHirschField", "hirsch_alpha", "DQMC", "run", "ExitCode",
           "LocalSweep", "GlobalFlip", "GlobalShuffle", "SpatialShuffle", "TemporalShuffle", "Denoise",
           "DenoiseFlip", "StaggeredDenoise", "SimpleScheduler", "AcceptanceStatistics",
           "ConfigRecorder", "Discarder", "save", "load", "resume", "MagnitudeStats", "DQMCAnalysis",
           "measurements"]
