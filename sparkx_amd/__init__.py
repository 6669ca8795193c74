"""sparkx_amd: MI355X-native HyperBall behind the reference's HyperBallInner API.

The engine is libsparkx_hb.so (HIP kernels for gfx950 + C ABI, include/sparkx_hb.h);
this package mirrors the reference's interface on the host."""
from .graph import Graph  # noqa: F401
from .hyper_ball import (HyperBall, HyperBall4, HyperBall5, HyperBall6, HyperBall7,  # noqa: F401
                         HyperBall8, HyperBall9, HyperBall10, HyperBall11, HyperBall12,
                         HyperBall13, HyperBall14, HyperBall15, HyperBall16)
from ._lib import HyperBallError  # noqa: F401
