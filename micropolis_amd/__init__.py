"""micropolis_amd: MI355X (gfx950) REYES hot path behind the micropolis
renderer interface (bound-and-split, dice, displacement, flat shading,
tiled rasterisation with a deterministic depth|id visibility buffer)."""
from .renderer import BEZIER, GREGORY, MicropolisError, Projection, Renderer  # noqa: F401
