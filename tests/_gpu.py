"""Shared helpers of the GPU tests (skip cleanly when no HIP device is present)."""

from __future__ import annotations

import pytest


def require_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from centrex_tlf_amd.lindblad.native import library, library_path

    if not library_path().exists():
        pytest.fail(f"{library_path()} is not built")
    library()
    return torch
