"""Koishi Radiance path tracer for MI355X (gfx950): Python bindings of the C ABI
(include/kr_api.h, include/koishi.h) and seeded scene generators."""
from . import kr, scenegen  # noqa: F401
