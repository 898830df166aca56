"""ising_amd -- MI355X-native Metropolis + parallel-tempering core for
classical Heisenberg spin lattices (the hot path of SamuelJosephs/ISING),
as hand-written gfx950 HIP kernels behind the C ABI of include/ising_gpu.h.
"""
from .engine import Engine, Hamiltonian, IsingError, Ladder, Lattice, geometric_temperatures

__all__ = ["Engine", "Hamiltonian", "IsingError", "Ladder", "Lattice", "geometric_temperatures"]
