"""The learned initial guess: a coordinate MLP u(x; punctures) on the library's BF16 MFMA
kernels (nrpyell_mlp_*), trained on solutions the solver itself produced.

Pipeline (nrpyelliptic_amd.learned drives it):
  features(problem)        -> device BF16 [rows x 64], per-row u scale, grid index
  targets(problem, ...)    -> device t = u/scale from a solved y_n
  InitialGuessMLP.train    -> Adam on uniform minibatches of the device dataset
  InitialGuessMLP.seed     -> y_n = network guess (u, v = eta u, ghost images)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .runtime import DeviceArray, check, lib

K0 = 64
WIDTH = 256


def _bind():
    L = lib()
    if getattr(L, "_mlp_bound", False):
        return L
    vp = C.c_void_p
    L.nrpyell_mlp_create.argtypes = [C.c_ulonglong, C.POINTER(vp)]
    L.nrpyell_mlp_destroy.argtypes = [vp]
    L.nrpyell_mlp_num_params.argtypes = [vp]
    L.nrpyell_mlp_num_params.restype = C.c_long
    L.nrpyell_mlp_get_params.argtypes = [vp, C.POINTER(C.c_float)]
    L.nrpyell_mlp_set_params.argtypes = [vp, C.POINTER(C.c_float)]
    L.nrpyell_mlp_forward.argtypes = [vp, vp, C.c_int, vp]
    L.nrpyell_mlp_loss_grad.argtypes = [vp, vp, vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_float)]
    L.nrpyell_mlp_adam_step.argtypes = [vp, C.c_float]
    L.nrpyell_mlp_train.argtypes = [vp, vp, vp, C.c_long, C.c_int, C.c_int, C.c_float, C.c_ulonglong,
                                    C.POINTER(C.c_double)]
    L.nrpyell_mlp_feature_rows.argtypes = [C.POINTER(abi.griddata_struct)]
    L.nrpyell_mlp_feature_rows.restype = C.c_long
    L.nrpyell_mlp_features.argtypes = [C.POINTER(abi.griddata_struct), vp, vp, vp]
    L.nrpyell_mlp_targets.argtypes = [C.POINTER(abi.griddata_struct), vp, vp, vp]
    L.nrpyell_mlp_seed.argtypes = [vp, C.POINTER(abi.griddata_struct), C.c_int]
    L.nrpyell_memcpy_d2d.argtypes = [vp, vp, C.c_size_t]
    L.nrpyell_mlp_sample_rows.argtypes = [vp, vp, C.c_long, C.c_long, C.c_ulonglong, vp, vp]
    L._mlp_bound = True
    return L


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float -> BF16 bit patterns (round to nearest even), as uint16."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = ((b >> 16) & 1) + np.uint32(0x7FFF)
    return ((b + r) >> 16).astype(np.uint16)


def bf16_to_float(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


class Features:
    """Device feature matrix of one problem: x (BF16 [rows x 64]), scale, idx; n valid rows."""

    def __init__(self, problem):
        L = _bind()
        self.rows = int(L.nrpyell_mlp_feature_rows(C.byref(problem.gd)))
        G0, N1, N2 = problem.Ng[0], problem.cfg.Nxx[1], problem.cfg.Nxx[2]
        self.n = (G0 - 5) * N1 * N2
        self.x = DeviceArray(self.rows * K0, np.uint16)
        self.scale = DeviceArray(self.rows, np.float64)
        self.idx = DeviceArray(self.rows, np.int64)
        check(L.nrpyell_mlp_features(C.byref(problem.gd), self.x.ptr, self.scale.ptr, self.idx.ptr), "features")

    def targets(self, problem) -> DeviceArray:
        t = DeviceArray(self.rows, np.float32).upload(np.zeros(self.rows, np.float32))
        check(_bind().nrpyell_mlp_targets(C.byref(problem.gd), self.scale.ptr, self.idx.ptr, t.ptr), "targets")
        return t

    def sample_into(self, t: DeviceArray, n_take: int, seed: int, x_dst: DeviceArray, t_dst: DeviceArray,
                    row0: int) -> None:
        """n_take random rows of (features, targets) into rows row0.. of a dataset."""
        xo = C.c_void_p(x_dst.ptr.value + row0 * K0 * 2)
        to = C.c_void_p(t_dst.ptr.value + row0 * 4)
        check(_bind().nrpyell_mlp_sample_rows(self.x.ptr, t.ptr, self.n, n_take, seed, xo, to), "sample_rows")


class InitialGuessMLP:
    def __init__(self, seed: int = 0):
        L = _bind()
        self.h = C.c_void_p()
        check(L.nrpyell_mlp_create(seed, C.byref(self.h)), "mlp_create")
        self.n_params = int(L.nrpyell_mlp_num_params(self.h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(_bind().nrpyell_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))), "get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(_bind().nrpyell_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "set_params")

    def save(self, path: str) -> None:
        np.savez_compressed(path, params=self.get_params(), k0=K0, width=WIDTH)

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as d:
            self.set_params(d["params"])

    def forward(self, x: DeviceArray, M: int) -> np.ndarray:
        y = DeviceArray(M, np.float32)
        check(_bind().nrpyell_mlp_forward(self.h, x.ptr, M, y.ptr), "mlp_forward")
        return y.download()

    def loss_grad(self, x: DeviceArray, t: DeviceArray, M: int):
        loss = C.c_double()
        g = np.empty(self.n_params, np.float32)
        check(_bind().nrpyell_mlp_loss_grad(self.h, x.ptr, t.ptr, M, C.byref(loss),
                                            g.ctypes.data_as(C.POINTER(C.c_float))), "loss_grad")
        return loss.value, g

    def train(self, x: DeviceArray, t: DeviceArray, N: int, batch: int, steps: int, lr: float, seed: int = 1):
        losses = np.zeros(steps, np.float64)
        check(_bind().nrpyell_mlp_train(self.h, x.ptr, t.ptr, N, batch, steps, lr, seed,
                                        losses.ctypes.data_as(C.POINTER(C.c_double))), "mlp_train")
        return losses

    def seed(self, problem, smooth_passes: int = 0) -> None:
        check(_bind().nrpyell_mlp_seed(self.h, C.byref(problem.gd), int(smooth_passes)), "mlp_seed")

    def close(self):
        if self.h:
            _bind().nrpyell_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
