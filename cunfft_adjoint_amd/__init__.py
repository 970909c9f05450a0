"""MI355X-native adjoint NFFT with the CUNA C API (libcunaf.so / libcunad.so).

The compute path is the C-ABI library built from csrc/ (HIP for gfx950);
this package is the Python mirror of its interface.  Build with
`python -m cunfft_adjoint_amd.build`.
"""
from ._lib import CunaError, load, lib_path  # noqa: F401
from .api import (  # noqa: F401
    CunaPlan,
    FilterProperties,
    Graph,
    autotune,
    capture,
    NFFTPlan,
    cunfft_adjoint,
    cunfft_adjoint_raw,
    cunfft_adjoint_raw_async,
    cunfft_adjoint_raw_async_bootstrap,
    cunfft_adjoint_raw_async_noalloc,
    cunfft_adjoint_raw_batched_async,
    batch_cache_clear,
    fft,
    generate_filter_properties,
    generate_pinned_filter_properties,
    generate_pinned_filter_properties_batch,
)

__version__ = "0.1.0"
