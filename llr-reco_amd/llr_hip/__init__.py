"""MI355X-native LLR diffusion-MRI reconstruction (Python front end).

Submodules mirror the Rust modules of wyatt-A/llr-reco whose API this library
keeps: block (block.rs), signal_model (signal_model.rs, cufft.rs),
dti_subspace (dti_subspace.rs), fftshift (fftshift.rs), low_rank (low_rank.rs,
cublas.rs) and recon (the llr_recon_exec loop of lib.rs). All compute runs in
libllr_hip.so through its C ABI (include/llr_hip.h); `phantom` builds seeded
synthetic inputs.
"""
from . import _lib
from ._lib import Context, LlrError, default_context, load

__all__ = ["Context", "LlrError", "default_context", "load", "block", "signal_model", "dti_subspace", "fftshift",
           "low_rank", "recon", "phantom"]


def __getattr__(name):
    if name in ("block", "signal_model", "dti_subspace", "fftshift", "low_rank", "recon", "phantom"):
        import importlib

        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
