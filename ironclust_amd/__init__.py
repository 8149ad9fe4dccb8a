"""ironclust_amd -- MI355X (gfx950) implementation of IronClust's density-peak hot path:
drift-constrained per-site kNN, rho = 1/d_k and delta/nneigh, behind a C ABI
(include/irc_hip.h, libirc_hip.so) and a host mirror of irc2's calling convention."""
from .libirc import Context, IrcError, load  # noqa: F401
from . import irc2, synth  # noqa: F401

__all__ = ["Context", "IrcError", "load", "irc2", "synth"]
