"""MI355X-native explicit element loop of OpenRadioss (8-node bricks, Belytschko-Tsay and
QEPH shells, law 2, PARITH/ON assembly, device nodal update) behind an OpenRadioss-style
C ABI (include/orad.h); engine.py binds it, autograd.py wraps the shell adjoint for torch."""
