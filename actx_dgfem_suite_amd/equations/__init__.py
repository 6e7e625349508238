"""Operator builders mirroring the reference's ``equations`` package."""
