"""Mirrors the reference package src/mpm_lbm/sim/monitoring."""
