"""lentils_amd: MI355X-native forward model and source inversion of Lentils.

Drop-in counterparts of lentils.common (spaces), lentils.models (lens
models) and lentils.operators (operator algebra, lens and Fourier
operators), backed by liblentils_hip.so (hand-written HIP kernels for
gfx950 + rocFFT), plus batched device-resident solvers.
"""
from .spaces import Space, VisibilitySpace, FourierSpace, ImageSpace, DelaunaySpace  # noqa: F401
from .models import GlobalLensModel, PowerLawEllipsoid, ExternalPotential  # noqa: F401
from .operators import (Operator, CompositeOperatorProduct, CompositeOperatorSum, DiagonalOperator,  # noqa: F401
                        PriorCovarianceOperator, ManifoldLensOperator, DelaunayLensOperator, FFTOperator, ZeroPaddingOperator,
                        ApodizationCorrectionOperator, GriddingOperator, NUFFTOperator, DFTOperator)
from .solvers import LensBatch, NUFFTPlan, BatchedSourceInversion  # noqa: F401
from .datasets import Dataset, RadioDataset, BlurredCovarianceOperator  # noqa: F401

__version__ = '0.1.0'
