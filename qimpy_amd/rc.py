"""Run configuration: device, streams and clock (reference: src/qimpy/rc.py).

One process drives one GPU. ``init()`` selects ``cuda:0`` (a ROCm device),
creates the secondary ``compute_stream`` at default priority (as
rc.py:119-128 does), and sets torch's CPU thread count to the physical cores
(rc.py:137-153). The GPU path has no CPU fallback: ``require_gpu()`` raises if
no ROCm device is visible.
"""
from __future__ import annotations

import os
import time
from typing import Optional

import torch

cpu = torch.device("cpu")
device: torch.device = cpu
use_cuda = False
compute_stream: Optional[torch.cuda.Stream] = None
t_start = time.time()
_initialized = False


def init(cores_override: Optional[int] = None) -> None:
    global device, use_cuda, compute_stream, t_start, _initialized
    t_start = time.time()
    if torch.cuda.is_available():
        device = torch.device("cuda:0")
        torch.cuda.set_device(device)
        use_cuda = True
        if os.environ.get("QIMPY_COMPUTE_STREAM", "1") in {"1", "yes"}:
            compute_stream = torch.cuda.Stream(device=device)  # default priority
    n_threads = cores_override or int(os.environ.get("SLURM_CPUS_PER_TASK", 0) or 0)
    if not n_threads:
        try:
            import psutil

            n_threads = psutil.cpu_count(logical=False) or os.cpu_count() or 1
        except ImportError:  # pragma: no cover
            n_threads = os.cpu_count() or 1
    torch.set_num_threads(n_threads)
    _initialized = True


def require_gpu() -> torch.device:
    """The GPU device; raises when none is visible (no CPU fallback)."""
    if not _initialized:
        init()
    if not use_cuda:
        raise RuntimeError("qimpy_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return device


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_get_device = getattr(torch._C, "_cuda_getDevice", None)


def stream_ptr(stream: Optional[torch.cuda.Stream] = None) -> int:
    """The hipStream_t of `stream`, by default torch's current stream on the
    current device (read without building a Stream object: every library
    call asks for it)."""
    if stream is not None:
        return int(stream.cuda_stream)
    if _raw_stream is not None and _get_device is not None:
        return int(_raw_stream(_get_device()))
    return int(torch.cuda.current_stream().cuda_stream)


def compute_stream_wait_current() -> None:
    if compute_stream is not None:
        compute_stream.wait_stream(torch.cuda.current_stream())


def current_stream_wait_compute() -> None:
    if compute_stream is not None:
        torch.cuda.current_stream().wait_stream(compute_stream)


def current_stream_synchronize() -> None:
    if use_cuda:
        torch.cuda.current_stream().synchronize()


def clock() -> float:
    return time.time() - t_start
