"""Auxiliary benchmark tools."""
