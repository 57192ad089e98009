"""nrpyelliptic_amd: MI355X (gfx950) hyperbolic-relaxation solver for NRPyElliptic puncture data."""
