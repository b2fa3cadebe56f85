"""Solver API and native boundary."""
from .solver import GridConfig, HipSimulator, Solver

__all__ = ["GridConfig", "HipSimulator", "Solver"]
