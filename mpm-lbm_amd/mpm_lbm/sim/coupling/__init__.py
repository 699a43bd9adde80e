"""Mirrors the reference package of the same name (src/mpm_lbm/sim/coupling)."""
