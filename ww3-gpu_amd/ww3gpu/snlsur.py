"""The learned W3SNL1 surrogate through the C ABI (ww3gpu_snlsur_*).

A multilayer perceptron (NSPEC + 1 inputs, three GELU layers of 512, NSPEC outputs; bf16 on
the matrix cores with fp32 accumulation) trained on this library's exact DIA, with an
energy-conservation check that sends failing points back to the exact DIA.  It has no
counterpart in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .capi import check, lib, ptr
from .snl1 import Snl1


class SnlSurrogate:
    """The surrogate of one configured DIA grid."""

    def __init__(self, dia: Snl1, seed: int = 0):
        self.dia = dia
        self.igrid = dia.igrid
        check("ww3gpu_snlsur_init", lib().ww3gpu_snlsur_init(self.igrid, seed))
        dims = np.zeros(4, np.int32)
        self.nparam = int(lib().ww3gpu_snlsur_nparams(self.igrid, ptr(dims)))
        self.kin, self.hidden, self.nout, self.nspec = (int(v) for v in dims)

    # parameters -----------------------------------------------------------------------
    def params(self) -> np.ndarray:
        p = np.zeros(self.nparam, np.float32)
        check("ww3gpu_snlsur_params", lib().ww3gpu_snlsur_params(self.igrid, 0, ptr(p), self.nparam))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check("ww3gpu_snlsur_params", lib().ww3gpu_snlsur_params(self.igrid, 1, ptr(p), p.size))

    def layers(self, p: np.ndarray = None) -> list:
        """[(W1, b1), ..., (W4, b4)] views of a parameter vector, W as [out][in]."""
        p = self.params() if p is None else p
        shapes = [(self.hidden, self.kin), (self.hidden, self.hidden), (self.hidden, self.hidden),
                  (self.nout, self.hidden)]
        out, o = [], 0
        for r, c in shapes:
            w = p[o:o + r * c].reshape(r, c)
            o += r * c
            b = p[o:o + r]
            o += r
            out.append((w, b))
        return out

    # data and training ---------------------------------------------------------------
    def synth(self, npts: int, seed: int, bimodal: float = 0.5):
        """Seeded training spectra on the device: (a, cg, kdmean) torch tensors."""
        import torch
        a = torch.empty((npts, self.nspec), dtype=torch.float32, device="cuda")
        cg = torch.empty((npts, self.dia.nk), dtype=torch.float32, device="cuda")
        kd = torch.empty((npts,), dtype=torch.float32, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        check("ww3gpu_snlsur_synth_dev",
              lib().ww3gpu_snlsur_synth_dev(self.igrid, npts, seed, bimodal, ptr(a), ptr(cg), ptr(kd), st))
        return a, cg, kd

    def fit(self, nsamples: int, batch: int, epochs: int, lr: float, seed: int = 1,
            bimodal: float = 0.5) -> np.ndarray:
        import torch
        losses = np.zeros(epochs, np.float32)
        st = torch.cuda.current_stream().cuda_stream
        check("ww3gpu_snlsur_fit", lib().ww3gpu_snlsur_fit(self.igrid, nsamples, batch, epochs, lr, seed, bimodal,
                                                           ptr(losses), st))
        return losses

    def train_step(self, a, cg, kd, s, lr: float) -> float:
        import torch
        loss = ctypes.c_float(0.0)
        st = torch.cuda.current_stream().cuda_stream
        check("ww3gpu_snlsur_train_dev",
              lib().ww3gpu_snlsur_train_dev(self.igrid, a.shape[0], ptr(a), ptr(cg), ptr(kd), ptr(s), lr,
                                            ctypes.addressof(loss), st))
        return float(loss.value)

    # inference -----------------------------------------------------------------------
    def predict(self, a, cg, kd):
        """S_nl for device tensors; returns (s, flags, nfallback) with flags[p] = 1 where the
        exact DIA was used."""
        import torch
        npts = a.shape[0]
        s = torch.empty_like(a)
        flags = torch.empty((npts,), dtype=torch.int32, device=a.device)
        nfb = ctypes.c_int(0)
        st = torch.cuda.current_stream().cuda_stream
        check("ww3gpu_snlsur_dev", lib().ww3gpu_snlsur_dev(self.igrid, npts, ptr(a), ptr(cg), ptr(kd), ptr(s),
                                                           ptr(flags), ctypes.addressof(nfb), st))
        return s, flags, int(nfb.value)

    # test entries --------------------------------------------------------------------
    def grads(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = np.zeros(x.shape[0], np.float32)
        g = np.zeros(self.nparam, np.float32)
        check("ww3gpu_snlsur_grads", lib().ww3gpu_snlsur_grads(self.igrid, x.shape[0], ptr(x), ptr(y), ptr(loss),
                                                               ptr(g)))
        return loss, g

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        o = np.zeros((x.shape[0], self.nout), np.float32)
        check("ww3gpu_snlsur_forward", lib().ww3gpu_snlsur_forward(self.igrid, x.shape[0], ptr(x), ptr(o)))
        return o


def gemm_nt(a, b):
    """C = A B^T on the matrix cores for bf16 torch tensors A (M, K), B (N, K)."""
    import torch
    M, K = a.shape
    N = b.shape[0]
    c = torch.empty((M, N), dtype=torch.float32, device=a.device)
    st = torch.cuda.current_stream().cuda_stream
    check("ww3gpu_sur_gemm_dev", lib().ww3gpu_sur_gemm_dev(M, N, K, ptr(a), K, ptr(b), K, ptr(c), N, st))
    return c


__all__ = ["SnlSurrogate", "gemm_nt"]
