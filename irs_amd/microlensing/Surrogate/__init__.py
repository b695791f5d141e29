"""Learned stand-in for full maps when only the magnification histogram is needed."""
