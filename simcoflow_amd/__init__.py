"""MI355X (gfx950) projection step and interface transport for
simcoflow-style 2D cut-cell two-phase flow."""
__all__ = ["cases", "projection"]
