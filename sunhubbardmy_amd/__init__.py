"""MI355X-native DQMC sweep (finite temperature), a drop-in for the sweep routines of the
Fortran package YuanyueChen/sunhubbardmy.  Compute runs in libsunhip.so (hand-written HIP for
gfx950); this package builds the model tables and drives the C ABI."""
from .model import Model, chain_seed, make_tables  # noqa: F401
from .walkers import FTWalkers, SweepError  # noqa: F401
from . import binio  # noqa: F401
