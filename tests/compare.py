"""Comparison helpers for HIP-vs-oracle parity tests."""
import torch


def rel_err(a, b):
    """max|a-b| / max(max|b|, tiny)"""
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    scale = max(b.abs().max().item(), 1e-300)
    return (a - b).abs().max().item() / scale


def assert_close(a, b, tol, what=""):
    e = rel_err(a, b)
    assert e <= tol, "%s: relative error %.3e > %.1e" % (what, e, tol)


def tol_for(dtype):
    # FP64: rounding-level agreement; FP32: the 1e-5 relative bar
    return 1e-11 if dtype == torch.float64 else 1e-5
