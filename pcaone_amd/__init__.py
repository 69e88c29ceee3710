"""MI355X-native randomized-SVD hot path of PCAone (C++/HIP library + thin ctypes binding)."""
from ._lib import PcaoneError, lib  # noqa: F401
