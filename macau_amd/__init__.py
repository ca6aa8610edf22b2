"""macau_amd -- MI355X-native Gibbs sampler for BPMF / Macau.

Drop-in for the sampling step of jaak-s/macau: the same Python entry points
(``macau``, ``bpmf``, ``make_train_test``, ``make_train_test_df``,
``MacauResult``, ``blockcg``) backed by hand-written HIP kernels for gfx950.
"""
__version__ = "0.1.0"

from .macau import macau, bpmf, MacauResult  # noqa: F401
from .macau import blockcg, make_train_test, make_train_test_df  # noqa: F401
