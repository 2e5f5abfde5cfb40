"""fsim: MI355X implementation of fairSIM's 2D SR-SIM reconstruction path.

Host side of libfsim.so: the vector API slot (``linalg``), the SIM
parameter / OTF mirror and reconstructions (``sim_algorithm``), the request
server (``livemode``) and the synthetic workload generator (``synthetic``).
"""
from . import _lib  # noqa: F401
