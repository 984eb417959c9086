"""MI355X-native boid-step engine (per-timestep update of BoidsSimulatorMPI).

The compute path is libboids_engine.so (hand-written gfx950 HIP kernels
behind the C ABI of include/boids_engine.h).  This package holds that
library, its sources (csrc/) and the ctypes mirror of the reference's
BoidSimulation class (engine.py).
"""
from .engine import (BoidSimulation, Engine, EngineError, build, generate_init, lib,  # noqa: F401
                     make_params)

__all__ = ["BoidSimulation", "Engine", "EngineError", "build", "generate_init", "lib", "make_params"]
