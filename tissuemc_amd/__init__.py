"""tissuemc_amd — MCML photon transport in layered tissue on the AMD MI355X.

The engine is HIP for gfx950 behind a C ABI (include/tissuemc_hip.h); this package binds it
(``engine``) and holds the analysis around it: batch statistics (``stats``), the semi-infinite
H-function reference (``hfunction``), published-value validation (``validate``) and the
identical-photon comparison (``cross_engine``).
"""

__version__ = "0.1.0"
