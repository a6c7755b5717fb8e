"""topaz_amd: MI355X (gfx950) implementation of TOPAZ's candidate-alignment
hot path -- gapless HSP extension and striped Smith-Waterman -- behind a C ABI
(include/topaz_amd.h) with this thin Python host layer."""
from ._lib import TopazError, lib  # noqa: F401
from .matrices import MATRIX_ORDER, matrix  # noqa: F401
