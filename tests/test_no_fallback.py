"""The product path has no CPU or non-HIP fallback: a missing libmpm_hip.so raises
ImportError, device="cpu" raises NotImplementedError, and without a visible GPU the HIP
Solver refuses to load particles instead of computing anywhere else (CPU)."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import Solver
from mpm_amd.core import native
from mpm_amd.core.solver import GridConfig

PTS = np.full((8, 3), 0.5, np.float32)
VOL = np.full(8, 1e-6, np.float32)


def test_missing_library_raises_import_error(monkeypatch, tmp_path):
    monkeypatch.setattr(native, "_LIB", None)
    monkeypatch.setattr(native, "LIB_PATH", str(tmp_path / "libmpm_hip.so"))
    with pytest.raises(ImportError, match="no non-HIP fallback"):
        native.load()


def test_binding_matches_the_header_abi_version():
    import os
    import re

    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "mpm_hip.h")).read()
    assert int(re.search(r"#define MPM_ABI_VERSION (\d+)", hdr).group(1)) == native.ABI_VERSION
    assert native.load().mpm_abi_version() == native.ABI_VERSION


def test_cpu_device_is_refused():
    with pytest.raises(NotImplementedError, match="only on the GPU"):
        Solver(grid=GridConfig(n_grid=32, grid_lim=1.0), device="cpu").load_particles(PTS, VOL)


def test_no_visible_gpu_is_refused():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible here")
    with pytest.raises(RuntimeError, match="no HIP device"):
        Solver(grid=GridConfig(n_grid=32, grid_lim=1.0), device="hip").load_particles(PTS, VOL)
