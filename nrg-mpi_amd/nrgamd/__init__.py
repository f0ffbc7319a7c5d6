"""Python helpers for the MI355X NRG hot path: input decks, the ctypes
binding of libnrg_gpu.so and the driver of the Fortran host library."""
