"""MI355X-native Krylov time evolution (TimeEvolver's hot path) — Python front-end.

The compute path is the C-ABI library libtimeevolver_hip.so built from csrc/ (hand-written
HIP for gfx950); this package only loads it through ctypes (see capi.py) and provides the
reference-style term factories (hamiltonian.py) and seeded workloads (workloads.py).
"""
