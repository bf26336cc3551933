"""MI355X (gfx950) implementation of CRAVA's Fourier-domain Bayesian AVO
inversion hot path: FFTGrid transforms, the per-wavenumber posterior and
posterior simulation, behind the C ABI in include/crava_hip.h."""
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libcrava_hip.so"

__all__ = ["PKG_DIR", "LIB_PATH"]
