"""MI355X-native MLP local search (host-side Python view of libwamca_mi.so)."""
from ._lib import LIB_PATH, MLPError, lib  # noqa: F401
from .engine import NB_NAMES, ROWMAP_ADS, ROWMAP_ODC, Engine, rows_of  # noqa: F401
