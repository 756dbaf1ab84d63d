"""Loads the native libraries of the package.

``torch`` is imported first: the HIP library links ``libamdhip64.so.7`` by soname and
must bind to the HIP runtime instance that torch has already loaded.  There is no
fallback: without the built extension every compute entry point fails loudly.
"""
from __future__ import annotations

import glob
import importlib.util
import os

import torch  # noqa: F401  (must be loaded before the native libraries)

_HERE = os.path.dirname(os.path.abspath(__file__))


def _load():
    cands = sorted(glob.glob(os.path.join(_HERE, "_fvdbk_ext*.so")))
    if not cands:
        raise ImportError(
            "fvdb-core_amd native extension not built: run `python fvdb-core_amd/build.py` "
            "(or __graft_entry__.build()) to compile the gfx950 HIP kernels"
        )
    spec = importlib.util.spec_from_file_location("fvdb._fvdbk_ext", cands[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


ext = _load()


def lib_path() -> str:
    return os.path.join(_HERE, "libfvdbk.so")


def require_device(t: torch.Tensor, what: str = "tensor") -> None:
    """The compute path runs on HIP devices only."""
    if not t.is_cuda:
        raise ValueError(f"{what} must be on a HIP device: fvdb-core_amd has no CPU compute path")
