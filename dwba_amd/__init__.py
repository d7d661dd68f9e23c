"""dwba_amd: MI355X-native radial multipole double integrals for DWBA.

I_L = int int rho_a(r1) r_<^L / r_>^(L+1) rho_b(r2) dr1 dr2, direct and
exchange, all L, on non-uniform radial grids, computed by hand-written gfx950
HIP kernels behind a C ABI (include/dwba_hip.h), called from Python through
ctypes.  ``dwba_amd.dwba_matrix_elements`` keeps the call signature of the
reference module of the same name.
"""
__version__ = "0.1.0"
