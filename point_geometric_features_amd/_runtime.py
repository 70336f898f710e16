"""HIP runtime selection for processes that also use PyTorch-ROCm.

PyTorch wheels bundle their own ``libamdhip64.so`` (same SONAME as the ROCm
install's).  If ``libpgeof_hip.so`` loads first it pulls in the ROCm install's
runtime and a later ``import torch`` loads a second, independent HIP runtime
in the same process, after which one of the two sees no GPU.  When PyTorch is
installed, its runtime is therefore loaded (by path, RTLD_GLOBAL) before this
package's library, so both share one runtime; otherwise the ROCm install's
runtime is used.  Importing torch itself is not needed.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
import sys

_done = False


def prefer_torch_hip_runtime() -> None:
    global _done
    if _done:
        return
    _done = True
    if "torch" in sys.modules:
        return  # torch already loaded its runtime; ours binds to it by SONAME
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
