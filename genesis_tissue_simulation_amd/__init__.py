"""MI355X-native XPBD tissue substep (endoscopic_tissue hot path) with hand-written HIP kernels for gfx950."""
__all__ = ["grid", "params", "tissue", "native_backend", "workloads"]
