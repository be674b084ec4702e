"""LocalKRRModel on the MI355X path: FCHL19 local-kernel KRR fit / predict / score / save / load.

Drop-in for the reference's high-level model (python/kernelforge/models/krr.py:24-324 and the
shared behaviour of models/base.py:55-424): the same constructor arguments, the same three
training modes inferred from which labels are given, the same per-element energy baseline,
the same ``.npz`` model keys (krr.py:286-324), so a model written by either side loads in the
other.  Every kernel, solve and descriptor here runs through this package's HIP library:

* descriptors and Jacobians: one batched launch (``kf_fchl19_f64``) for the whole set,
  instead of one call per molecule (models/representations.py:102-150);
* training kernels: ``kernel_gaussian_symm_rfp`` / ``_hessian_symm_rfp`` /
  ``_full_symm_rfp`` straight into RFP, solved by ``cho_solve_rfp``;
* prediction: the rectangular and Jacobian kernels (energy mode) or the matrix-free
  descriptor-space matvecs with the precomputed alpha_desc (force and E+F modes).

The model is host-side orchestration; it holds NumPy arrays, as the reference's does.
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

from . import capi, kernelmath, local_kernels

ENERGY_ONLY, FORCE_ONLY, ENERGY_AND_FORCE = "energy_only", "force_only", "energy_and_force"
DEFAULT_ELEMENTS = (1, 6, 7, 8, 16)
# keyword arguments of generate_fchl_acsf(_and_gradients) a model may carry in repr_params
REPR_KEYS = ("nRs2", "nRs3", "nFourier", "eta2", "eta3", "zeta", "rcut", "acut",
             "two_body_decay", "three_body_decay", "three_body_weight")


@dataclass
class ModelScore:
    """Regression statistics of one predicted quantity (models/base.py:13-39)."""

    mae: float
    pearson_r: float
    slope: float
    intercept: float

    def __str__(self) -> str:
        return (f"MAE={self.mae:.4f}  r={self.pearson_r:.6f}  "
                f"slope={self.slope:.6f}  intercept={self.intercept:.4f}")


def score_arrays(y_ref, y_pred) -> ModelScore:
    """MAE, Pearson r and the least-squares line y_pred ~ slope * y_ref + intercept."""
    ref = np.asarray(y_ref, dtype=np.float64).ravel()
    pred = np.asarray(y_pred, dtype=np.float64).ravel()
    slope, intercept = np.polyfit(ref, pred, 1)
    return ModelScore(mae=float(np.mean(np.abs(pred - ref))),
                      pearson_r=float(np.corrcoef(ref, pred)[0, 1]),
                      slope=float(slope), intercept=float(intercept))


def training_mode(energies, forces) -> str:
    """energies only -> energy_only, forces only -> force_only, both -> energy_and_force."""
    if energies is None and forces is None:
        raise ValueError("At least one of `energies` or `forces` must be provided.")
    if forces is None:
        return ENERGY_ONLY
    return FORCE_ONLY if energies is None else ENERGY_AND_FORCE


def as_geometry(coords_list, z_list):
    """Per-molecule float64 coordinates and int32 charges; the two lists must match."""
    if len(coords_list) != len(z_list):
        raise ValueError(f"coords_list and z_list must have the same length, got "
                         f"{len(coords_list)} and {len(z_list)}")
    return ([np.asarray(c, dtype=np.float64) for c in coords_list],
            [np.asarray(z, dtype=np.int32) for z in z_list])


def as_forces(forces) -> np.ndarray:
    """Forces as (n_mols, 3 n_atoms) when every molecule has the same size, else flat.

    Accepts a stacked (n_mols, n_atoms, 3) / (n_mols, 3 n_atoms) array or a list of
    per-molecule (n_i, 3) / (3 n_i,) arrays (models/base.py:352-372)."""
    if isinstance(forces, np.ndarray):
        f = np.asarray(forces, dtype=np.float64)
        if f.ndim == 3:
            return f.reshape(f.shape[0], -1)
        if f.ndim == 2:
            return f
        raise ValueError(f"forces ndarray must be 2D or 3D, got {f.ndim}D")
    rows = [np.asarray(fi, dtype=np.float64).ravel() for fi in forces]
    if len({r.size for r in rows}) == 1:
        return np.stack(rows)
    return np.concatenate(rows)


def composition(z_list, elements) -> np.ndarray:
    """A[i, j] = atoms of charge elements[j] in molecule i."""
    el = np.asarray(elements, dtype=np.int32)
    A = np.zeros((len(z_list), el.size))
    for i, z in enumerate(z_list):
        A[i] = (np.asarray(z)[:, None] == el[None, :]).sum(axis=0)
    return A


def element_baseline(z_list, energies):
    """Per-element energies e minimising |A e - E| (no intercept) over the elements present."""
    elements = np.unique(np.concatenate([np.asarray(z, dtype=np.int32) for z in z_list]))
    elements = elements.astype(np.int32)
    e, *_ = np.linalg.lstsq(composition(z_list, elements), np.asarray(energies, np.float64),
                            rcond=None)
    return elements, np.asarray(e, dtype=np.float64)


def compute_fchl19(coords_list, z_list, elements, with_gradients=True, repr_params=None):
    """FCHL19 descriptors of a molecule list in one batched GPU launch.

    Returns X (n_mols, max_atoms, rep), dX (n_mols, max_atoms, rep, 3 max_atoms) or None,
    Q (n_mols, max_atoms) nuclear charges and N (n_mols,), zero-padded past N[i], as the
    reference's models/representations.py:102-150 lays them out.  With gradients the
    three-body term keeps the first angular harmonic only, like the reference's gradient
    path (src/fchl19_repr.cpp:449,523-528)."""
    params = dict(repr_params or {})
    unknown = sorted(set(params) - set(REPR_KEYS))
    if unknown:
        raise ValueError(f"unknown FCHL19 parameters {unknown}")
    coords_list, z_list = as_geometry(coords_list, z_list)
    nm = len(z_list)
    N = np.array([z.size for z in z_list], dtype=np.int32)
    ma = int(N.max()) if nm else 0
    C = np.zeros((nm, ma, 3))
    Q = np.zeros((nm, ma), dtype=np.int32)
    for i, (c, z) in enumerate(zip(coords_list, z_list)):
        C[i, : z.size] = np.asarray(c).reshape(z.size, 3)
        Q[i, : z.size] = z
    nfourier = int(params.get("nFourier", 1))
    nab = 2 * nfourier
    if with_gradients:
        nab = max(nab, 2)
    rep = capi.rep_size(len(elements), int(params.get("nRs2", 24)), int(params.get("nRs3", 20)),
                        nab)
    DA = capi.DeviceArray
    dC, dQ, dN = DA.from_numpy(C), DA.from_numpy(Q), DA.from_numpy(N)
    X = DA((nm, ma, rep))
    dX = DA((nm, ma, rep, 3 * ma)) if with_gradients else None
    capi.fchl19(dC.ptr, dQ.ptr, dN.ptr, nm, ma, X.ptr, dX.ptr if dX is not None else 0,
                elements=tuple(int(e) for e in elements), **params)
    capi.synchronize()
    return X.numpy(), (dX.numpy() if dX is not None else None), Q, N


class LocalKRRModel:
    """Kernel ridge regression on FCHL19 local descriptors (models/krr.py:24-66).

    Parameters
    ----------
    sigma : Gaussian length-scale.
    l2 : regularisation added to the kernel diagonal.
    elements : sorted atomic numbers of the descriptor channels (default H, C, N, O, S).
    repr_params : extra FCHL19 keyword arguments (nRs2, nRs3, rcut, eta2, ...).
    """

    def __init__(self, sigma: float = 2.0, l2: float = 1e-9, elements=None,
                 repr_params: dict[str, Any] | None = None) -> None:
        self.sigma = float(sigma)
        self.l2 = float(l2)
        self.elements = sorted(int(e) for e in (elements if elements is not None
                                                else DEFAULT_ELEMENTS))
        self.repr_params = dict(repr_params) if repr_params is not None else {}
        self.is_fitted_ = False

    # ------------------------------------------------------------------ public interface
    def fit(self, coords_list, z_list, energies=None, forces=None) -> "LocalKRRModel":
        """Train on energies, forces (F = -dE/dR) or both (mode inferred from which are given).

        With energies, a per-element baseline is fitted by least squares and subtracted
        before training; predict() adds it back."""
        coords_list, z_list = as_geometry(coords_list, z_list)
        self.training_mode_ = training_mode(energies, forces)
        E = None if energies is None else np.asarray(energies, dtype=np.float64).ravel()
        F = None if forces is None else as_forces(forces)
        if E is not None:
            self.baseline_elements_, self.element_energies_ = element_baseline(z_list, E)
            E = E - composition(z_list, self.baseline_elements_) @ self.element_energies_
        else:
            self.baseline_elements_ = np.array([], dtype=np.int32)
            self.element_energies_ = np.array([], dtype=np.float64)
        self._fit(coords_list, z_list, E, F)
        self.train_score_ = {k: score_arrays(ref, pred)
                             for k, (ref, pred) in self._training_fit().items()}
        self.is_fitted_ = True
        return self

    def predict(self, coords_list, z_list, compute_energy: bool = True):
        """(energies (M,), forces) for new structures.  Forces are F = -dE/dR as one flat
        array of 3 * sum(n_atoms) values, molecule by molecule (the layout the reference's
        kernels return).  With compute_energy=False the energies are zeros."""
        self._check_fitted()
        coords_list, z_list = as_geometry(coords_list, z_list)
        E, F = self._predict(coords_list, z_list)
        if not compute_energy:
            E = np.zeros(len(z_list))
        elif self.baseline_elements_.size:
            E = E + composition(z_list, self.baseline_elements_) @ self.element_energies_
        return E, F

    def score(self, coords_list, z_list, energies=None, forces=None) -> dict[str, ModelScore]:
        """ModelScore per quantity given ("energy", "force") against predict()."""
        if energies is None and forces is None:
            raise ValueError("At least one of `energies` or `forces` must be provided to score().")
        E_pred, F_pred = self.predict(coords_list, z_list)
        out = {}
        if energies is not None:
            out["energy"] = score_arrays(energies, E_pred)
        if forces is not None:
            if self.training_mode_ == ENERGY_ONLY:
                raise ValueError("Cannot score forces: model was trained in energy_only mode "
                                 "and does not predict forces.")
            out["force"] = score_arrays(as_forces(forces), F_pred)
        return out

    def save(self, path) -> None:
        """Write the model to ``path`` (.npz; the extension is added when missing)."""
        self._check_fitted()
        path = Path(path)
        if path.suffix != ".npz":
            path = path.with_suffix(".npz")
        np.savez(path, **self._arrays())

    @classmethod
    def load(cls, path) -> "LocalKRRModel":
        """Read a model written by save() (or by the reference's LocalKRRModel.save)."""
        model = cls.__new__(cls)
        with np.load(Path(path), allow_pickle=False) as data:
            model._from_arrays(data)
        model.is_fitted_ = True
        return model

    # ------------------------------------------------------------------ internals
    def _check_fitted(self):
        if not getattr(self, "is_fitted_", False):
            raise RuntimeError(f"{type(self).__name__} is not fitted yet. Call fit() first.")

    def _descriptors(self, coords_list, z_list):
        return compute_fchl19(coords_list, z_list, self.elements, True, self.repr_params)

    def _fit(self, coords_list, z_list, E, F):
        X, dX, Q, N = self._descriptors(coords_list, z_list)
        self._X_tr, self._dX_tr, self._Q_tr, self._N_tr = X, dX, Q, N
        self._n_train = len(z_list)
        s = self.sigma
        if self.training_mode_ == ENERGY_ONLY:
            y = E
            K = local_kernels.kernel_gaussian_symm_rfp(X, Q, N, s)
        elif self.training_mode_ == FORCE_ONLY:
            y = F.ravel()
            K = local_kernels.kernel_gaussian_hessian_symm_rfp(X, dX, Q, N, s)
        else:
            # the full kernel's force rows are +dK/dR, so the labels are [E; -F] = [E; dE/dR]
            y = np.concatenate([E, -F.ravel()])
            K = local_kernels.kernel_gaussian_full_symm_rfp(X, dX, Q, N, s)
        if K.size != y.size * (y.size + 1) // 2:
            raise ValueError(f"{y.size} labels for a kernel of order "
                             f"{int((np.sqrt(8 * K.size + 1) - 1) // 2)}")
        self._y_train = y
        self._alpha = kernelmath.cho_solve_rfp(K, y, l2=self.l2)
        self._alpha_desc = None
        if self.training_mode_ != ENERGY_ONLY:
            alpha_F = self._alpha if self.training_mode_ == FORCE_ONLY else self._alpha[self._n_train:]
            self._alpha_desc = local_kernels.kernel_gaussian_local_compute_alpha_desc(
                dX, Q, N, alpha_F)

    def _predict(self, coords_list, z_list):
        Xt, dXt, Qt, Nt = self._descriptors(coords_list, z_list)
        X, dX, Q, N, a, s = self._X_tr, self._dX_tr, self._Q_tr, self._N_tr, self._alpha, self.sigma
        mode = self.training_mode_
        if mode == ENERGY_ONLY:
            E = local_kernels.kernel_gaussian(Xt, X, Qt, Q, Nt, N, s) @ a
            F = local_kernels.kernel_gaussian_jacobian_t(Xt, dXt, X, Qt, Q, Nt, N, s) @ a
        elif mode == FORCE_ONLY:
            if self._alpha_desc is not None:
                F = local_kernels.kernel_gaussian_local_hessian_matvec(
                    Xt, dXt, X, self._alpha_desc, Qt, Q, Nt, N, s)
                E = -local_kernels.kernel_gaussian_local_jacobian_t_matvec(
                    Xt, X, self._alpha_desc, Qt, Q, Nt, N, s)
            else:  # a model file without alpha_desc: explicit kernels
                F = local_kernels.kernel_gaussian_hessian(Xt, dXt, X, dX, Qt, Q, Nt, N, s) @ a
                E = -(local_kernels.kernel_gaussian_jacobian(Xt, X, dX, Qt, Q, Nt, N, s) @ a)
        else:
            if self._alpha_desc is not None:
                E, Fneg = local_kernels.kernel_gaussian_local_full_matvec(
                    Xt, dXt, X, self._alpha_desc, a[: self._n_train], Qt, Q, Nt, N, s)
                F = -Fneg
            else:
                y = local_kernels.kernel_gaussian_full(Xt, dXt, X, dX, Qt, Q, Nt, N, s) @ a
                E, F = y[: len(z_list)], -y[len(z_list):]
        return np.asarray(E, dtype=np.float64), np.asarray(F, dtype=np.float64)

    def _training_fit(self):
        """(labels, fitted values) on the training set: (K + l2 I) alpha = y gives
        K alpha = y - l2 alpha, so no kernel is evaluated."""
        pred = self._y_train - self.l2 * self._alpha
        if self.training_mode_ == ENERGY_ONLY:
            return {"energy": (self._y_train, pred)}
        if self.training_mode_ == FORCE_ONLY:
            return {"force": (self._y_train, pred)}
        n = self._n_train
        return {"energy": (self._y_train[:n], pred[:n]),
                "force": (-self._y_train[n:], -pred[n:])}

    def _arrays(self) -> dict[str, object]:
        """The reference's model keys (models/krr.py:286-307)."""
        out: dict[str, object] = {
            "model_class": "LocalKRRModel",
            "training_mode": self.training_mode_,
            "sigma": self.sigma,
            "l2": self.l2,
            "elements": np.array(self.elements, dtype=np.int32),
            "repr_params": json.dumps(self.repr_params),
            "baseline_elements": self.baseline_elements_,
            "element_energies": self.element_energies_,
            "alpha": self._alpha,
            "y_train": self._y_train,
            "X_tr": self._X_tr,
            "Q_tr": self._Q_tr,
            "N_tr": self._N_tr,
        }
        if self._dX_tr is not None:
            out["dX_tr"] = self._dX_tr
        if self._alpha_desc is not None:
            out["alpha_desc"] = self._alpha_desc
        return out

    def _from_arrays(self, data) -> None:
        if "model_class" in data and str(data["model_class"]) != "LocalKRRModel":
            raise ValueError(f"not a LocalKRRModel file (model_class={data['model_class']})")
        self.sigma = float(data["sigma"])
        self.l2 = float(data["l2"])
        self.elements = [int(e) for e in data["elements"]]
        self.repr_params = json.loads(str(data["repr_params"]))
        self.baseline_elements_ = np.asarray(data["baseline_elements"], dtype=np.int32)
        self.element_energies_ = np.asarray(data["element_energies"], dtype=np.float64)
        self.training_mode_ = str(data["training_mode"])
        self._alpha = np.asarray(data["alpha"], dtype=np.float64)
        self._y_train = (np.asarray(data["y_train"], dtype=np.float64) if "y_train" in data
                         else np.array([], dtype=np.float64))
        self._X_tr = np.asarray(data["X_tr"], dtype=np.float64)
        self._Q_tr = np.asarray(data["Q_tr"], dtype=np.int32)
        self._N_tr = np.asarray(data["N_tr"], dtype=np.int32)
        self._dX_tr = np.asarray(data["dX_tr"], dtype=np.float64) if "dX_tr" in data else None
        self._alpha_desc = (np.asarray(data["alpha_desc"], dtype=np.float64)
                            if "alpha_desc" in data else None)
        self._n_train = int(self._N_tr.size)
