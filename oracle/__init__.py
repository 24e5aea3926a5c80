"""FP64 CPU oracle of the fftgal power-spectrum path (test infrastructure).

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
package; it is the checker the GPU library is compared against, never part of
the product path.  The C restatement lives in oracle.c / fft64.c and is built
into oracle/liboracle.so by oracle/Makefile.
"""
from .oracle import (  # noqa: F401
    load, Grid, Catalog, hat, NK, power_pipeline, fft_r2c_3d, fft_c2r_3d, set_threads,
    default_threads,
)
