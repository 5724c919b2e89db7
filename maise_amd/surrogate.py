"""The screening surrogate: a nin-64-64-1 tanh network trained and evaluated by the library's own HIP kernels
(mg_sur_* in libmaise_gpu.so).

It maps per-structure descriptor features of an unrelaxed structure (Context.features: species means of the
same-species descriptor block) to the energy per atom that structure reaches after relaxation.  Training is
mini-batch Adam on the mean-squared error with inputs and target standardised from the training data.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _fp, lib


class Surrogate:
    H = 64

    def __init__(self, nin: int, device: int = 0, seed: int = 1):
        self.nin = int(nin)
        self.P = int(lib().mg_sur_nparams(self.nin))
        self._h = C.c_void_p()
        _check(lib().mg_sur_create(int(device), self.nin, int(seed), C.byref(self._h)), "mg_sur_create")
        self.norm = (np.zeros(self.nin, np.float32), np.ones(self.nin, np.float32), 0.0, 1.0)

    def close(self):
        if self._h:
            lib().mg_sur_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.P, dtype=np.float32)
        _check(lib().mg_sur_get_params(self._h, _fp(p)), "mg_sur_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.P:
            raise ValueError(f"expected {self.P} parameters, got {p.size}")
        _check(lib().mg_sur_set_params(self._h, _fp(p)), "mg_sur_set_params")

    def set_norm(self, mu, sig, ymu: float, ysig: float):
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        sig = np.ascontiguousarray(sig, dtype=np.float32)
        _check(lib().mg_sur_set_norm(self._h, _fp(mu), _fp(sig), float(ymu), float(ysig)), "mg_sur_set_norm")
        self.norm = (mu, sig, float(ymu), float(ysig))

    def fit_norm(self, X, y, floor: float = 1e-6):
        """Standardise inputs and target with the mean and standard deviation of the training data (a constant
        column keeps unit scale)."""
        X = np.asarray(X, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        sig = X.std(axis=0)
        sig = np.where(sig > floor, sig, 1.0)
        ys = float(y.std()) if y.std() > floor else 1.0
        self.set_norm(X.mean(axis=0), sig, float(y.mean()), ys)

    def train(self, X, y, epochs: int = 20, batch: int = 64, lr: float = 1e-3, seed: int = 1) -> float:
        """Mini-batch Adam; returns the final mean-squared error over (X, y) in standardised units."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.nin)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        loss = C.c_float()
        _check(lib().mg_sur_train(self._h, _fp(X), _fp(y), len(y), int(epochs), int(batch), float(lr), int(seed),
                                  C.byref(loss)), "mg_sur_train")
        return loss.value

    def train_begin(self, X, y, epochs: int = 20, batch: int = 64, lr: float = 1e-3, seed: int = 1) -> None:
        """The same training run left in flight on the device; train_end() waits for it and returns the loss."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.nin)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        _check(lib().mg_sur_train_begin(self._h, _fp(X), _fp(y), len(y), int(epochs), int(batch), float(lr),
                                        int(seed)), "mg_sur_train_begin")

    def train_end(self) -> float:
        loss = C.c_float()
        _check(lib().mg_sur_train_end(self._h, C.byref(loss)), "mg_sur_train_end")
        return loss.value

    def grad(self, X, y):
        """(loss, gradient) of the mean-squared error over all of (X, y); parameters unchanged."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.nin)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        g = np.zeros(self.P, dtype=np.float32)
        loss = C.c_float()
        _check(lib().mg_sur_grad(self._h, _fp(X), _fp(y), len(y), _fp(g), C.byref(loss)), "mg_sur_grad")
        return loss.value, g

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.nin)
        out = np.zeros(len(X), dtype=np.float32)
        _check(lib().mg_sur_predict(self._h, _fp(X), len(X), _fp(out)), "mg_sur_predict")
        return out
