"""Simulation events (the reference's events package, events/__init__.py:1-11):
the state an event sees, the event, and the iteration and time schedulers.
They are defined with the simulation loop that runs them (simulation.py)."""

from ..simulation import Event, IterationScheduler, SimulationState, TimeScheduler

__all__ = ["Event", "SimulationState", "IterationScheduler", "TimeScheduler"]
