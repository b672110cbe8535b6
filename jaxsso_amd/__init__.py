"""jaxsso_amd -- the JaxSSO finite-element forward + adjoint path on MI355X (gfx950).

Modules mirror the reference's (``JaxSSO/{model,SSO_model,assemblemodel,solver,element}.py``):

    import jaxsso_amd.model as Model
    from jaxsso_amd.SSO_model import NodeParameter, ElementParameter, SSO_model

Compute runs in ``_lib/libjsso_hip.so`` (hand-written HIP kernels, C ABI in
``include/jsso.h``); PyTorch provides device memory, streams and autograd.
"""
from . import workloads  # noqa: F401
from ._native import NativeUnavailable, JssoError  # noqa: F401

__all__ = ["model", "SSO_model", "assemblemodel", "solver", "element", "plan", "batch", "autograd", "workloads"]
__version__ = "0.1.0"
