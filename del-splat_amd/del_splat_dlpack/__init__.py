"""del_splat_dlpack for MI355X: the reference's Python package
(del-splat-dlpack/del_splat_dlpack) over the dsplat HIP library.

Sub-packages keep the reference's names and shapes: Points, SplatCircle,
SplatPoint (del_splat_dlpack/{Points,SplatCircle,SplatPoint}), plus
SplatGauss (the Gaussian pipeline and the differentiable renderer).
"""


class _CapsuleAsDLPack:
    """Wraps a DLPack capsule for torch.from_dlpack (__init__.py:1-14 of the
    reference package); the device is read from the capsule itself."""

    def __init__(self, cap, device=(1, 0)):
        self._cap = cap
        self._used = False
        self._device = device

    def __dlpack__(self, *, max_version=None, stream=None, dl_device=None, copy=None):
        if self._used:
            raise RuntimeError("DLPack capsule already consumed")
        self._used = True
        return self._cap

    def __dlpack_device__(self):
        return self._device
