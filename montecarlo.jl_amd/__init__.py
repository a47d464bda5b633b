from .lattice import Chain, SquareLattice, CubicLattice, Honeycomb, TriangularLattice, Lattice
from .model import HubbardModel, AttractiveHubbardModel, RepulsiveHubbardModel
from .parameters import DQMCParameters, generate_chunks
from .fields import DensityHirschField, MagneticHirschField
