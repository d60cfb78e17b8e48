"""pdes_amd: MI355X-native Euler residual and Newton-Krylov hot path of
PDESolver.jl, behind a C ABI (include/pdes.h)."""
