"""Integrations of the GPU MVP plans with external solvers."""
