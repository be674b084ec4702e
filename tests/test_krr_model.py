"""LocalKRRModel host logic (no GPU): training-mode inference, label coercion, the
per-element energy baseline, scores, and the .npz model file (the reference's keys,
models/krr.py:286-307; readable without unpickling)."""

import numpy as np
import pytest

from kernelforge_amd import krr

REFERENCE_KEYS = {"model_class", "training_mode", "sigma", "l2", "elements", "repr_params",
                  "baseline_elements", "element_energies", "alpha", "y_train", "X_tr", "Q_tr",
                  "N_tr", "dX_tr", "alpha_desc"}


def test_training_mode():
    assert krr.training_mode(np.zeros(2), None) == "energy_only"
    assert krr.training_mode(None, np.zeros((2, 6))) == "force_only"
    assert krr.training_mode(np.zeros(2), np.zeros((2, 6))) == "energy_and_force"
    with pytest.raises(ValueError):
        krr.training_mode(None, None)


def test_as_forces_shapes():
    f3 = np.arange(24.0).reshape(2, 4, 3)
    assert krr.as_forces(f3).shape == (2, 12)
    assert krr.as_forces(f3.reshape(2, 12)).shape == (2, 12)
    assert krr.as_forces([f3[0], f3[1]]).shape == (2, 12)
    ragged = krr.as_forces([np.ones((2, 3)), np.ones((3, 3))])
    assert ragged.shape == (15,)
    with pytest.raises(ValueError):
        krr.as_forces(np.zeros(5))


def test_composition_and_baseline_recovers_element_energies():
    rng = np.random.default_rng(0)
    z_list = [rng.choice([1, 6, 8], size=rng.integers(3, 9)).astype(np.int32) for _ in range(40)]
    A = krr.composition(z_list, [1, 6, 8])
    assert A.shape == (40, 3)
    assert np.all(A.sum(axis=1) == [z.size for z in z_list])
    e_true = np.array([-0.5, -38.0, -75.0])
    els, e = krr.element_baseline(z_list, A @ e_true)
    np.testing.assert_array_equal(els, [1, 6, 8])
    np.testing.assert_allclose(e, e_true, rtol=1e-12, atol=1e-10)


def test_score_arrays():
    ref = np.linspace(-1.0, 2.0, 11)
    s = krr.score_arrays(ref, 2.0 * ref + 0.5)
    assert s.pearson_r == pytest.approx(1.0)
    assert s.slope == pytest.approx(2.0) and s.intercept == pytest.approx(0.5)
    assert s.mae == pytest.approx(np.mean(np.abs(ref + 0.5)))
    assert "MAE=" in str(s)


def _fake_model(mode):
    m = krr.LocalKRRModel(sigma=1.5, l2=1e-7, elements=[8, 1, 6], repr_params={"nRs2": 12})
    rng = np.random.default_rng(1)
    m.training_mode_ = mode
    m.baseline_elements_ = np.array([1, 6, 8], dtype=np.int32)
    m.element_energies_ = rng.standard_normal(3)
    m._n_train = 3
    m._X_tr = rng.standard_normal((3, 4, 10))
    m._Q_tr = np.array([[1, 6, 8, 0]] * 3, dtype=np.int32)
    m._N_tr = np.array([3, 3, 3], dtype=np.int32)
    m._dX_tr = rng.standard_normal((3, 4, 10, 12))
    m._alpha = rng.standard_normal(3 + 27)
    m._y_train = rng.standard_normal(3 + 27)
    m._alpha_desc = rng.standard_normal((3, 4, 10))
    m.is_fitted_ = True
    return m


def test_save_load_roundtrip_reference_keys(tmp_path):
    m = _fake_model("energy_and_force")
    m.save(tmp_path / "model")  # .npz added
    path = tmp_path / "model.npz"
    with np.load(path, allow_pickle=False) as data:
        assert set(data.files) == REFERENCE_KEYS
        assert str(data["model_class"]) == "LocalKRRModel"
    m2 = krr.LocalKRRModel.load(path)
    assert m2.is_fitted_ and m2.training_mode_ == "energy_and_force"
    assert (m2.sigma, m2.l2, m2.elements, m2.repr_params) == (1.5, 1e-7, [1, 6, 8], {"nRs2": 12})
    for k in ("_alpha", "_y_train", "_X_tr", "_Q_tr", "_N_tr", "_dX_tr", "_alpha_desc",
              "baseline_elements_", "element_energies_"):
        np.testing.assert_array_equal(getattr(m2, k), getattr(m, k))
    assert m2._n_train == 3
    # training fit without kernels: y - l2 alpha, forces with the sign of the labels undone
    fit = m2._training_fit()
    np.testing.assert_allclose(fit["force"][1], -(m._y_train[3:] - 1e-7 * m._alpha[3:]))


def test_load_rejects_other_model_classes(tmp_path):
    m = _fake_model("energy_only")
    arrays = m._arrays()
    arrays["model_class"] = "GlobalKRRModel"
    np.savez(tmp_path / "other.npz", **arrays)
    with pytest.raises(ValueError):
        krr.LocalKRRModel.load(tmp_path / "other.npz")


def test_unfitted_and_bad_inputs():
    m = krr.LocalKRRModel()
    with pytest.raises(RuntimeError):
        m.predict([np.zeros((2, 3))], [np.array([1, 1])])
    with pytest.raises(ValueError):
        krr.as_geometry([np.zeros((2, 3))], [])
    with pytest.raises(ValueError):
        krr.compute_fchl19([np.zeros((2, 3))], [np.array([1, 1])], [1], repr_params={"bogus": 1})
    with pytest.raises(ValueError):
        m.score([np.zeros((2, 3))], [np.array([1, 1])])
