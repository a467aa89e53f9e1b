"""Preset compositions per ocean model (mirrors offline_particles.models)."""
