"""MI355X-native VlaPy core: the 1D-1V Vlasov-Poisson-Fokker-Planck step on
hand-written HIP kernels (gfx950) behind the vlapy.core call surface."""

__version__ = "0.1.0"
