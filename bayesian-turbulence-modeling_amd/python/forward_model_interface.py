"""Black-box forward-model protocol on the MI355X engine.

Mirrors the boundary part of the reference's python/forward_model_interface.py
(:42-155): EvaluationResult, ForwardModelBase and the wrappers of the C++
incompressible and compressible forward models, with the same evaluate(theta) -> EvaluationResult
(.predictions list[float], .status str) contract that BayesianInferenceKOH
consumes.  EnsembleForwardModel adds the batched form: many thetas per call,
solved concurrently on the device.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class EvaluationResult:
    predictions: List[float]
    converged: bool = True
    status: str = "ok"
    metadata: Dict = field(default_factory=dict)

    def __bool__(self) -> bool:
        return self.converged and bool(self.predictions)


class ForwardModelBase(ABC):
    @abstractmethod
    def evaluate(self, theta) -> EvaluationResult:
        """Evaluate the forward model at parameter vector ``theta``."""

    @abstractmethod
    def parameter_names(self) -> List[str]:
        """Names of the input parameters."""

    def precomputed_ensemble(self) -> Optional[Tuple[np.ndarray, np.ndarray]]:
        return None


_META = (("simple_iters", "simple_iters"), ("log_lik", "log_lik"), ("latency_s", "latency"))


def _wrap(result) -> EvaluationResult:
    """Reference python/forward_model_interface.py:116-124: predictions,
    converged = non-empty and all finite, status as text.  Only .predictions
    and .status are required of the C++ result; the engine's iteration count,
    log-likelihood and latency go to metadata where the result has them."""
    preds = list(result.predictions) if result.predictions else []
    converged = bool(preds) and all(np.isfinite(p) for p in preds)
    meta = {key: getattr(result, attr) for key, attr in _META if hasattr(result, attr)}
    return EvaluationResult(predictions=preds, converged=converged, status=str(result.status), metadata=meta)


def _param_names(param_set) -> List[str]:
    """Reference python/forward_model_interface.py:541-547: the set's own
    names, a dict's "names", else theta_0 .. theta_{n-1} (n = n_active(),
    default 2)."""
    if hasattr(param_set, "active_names"):
        return list(param_set.active_names())
    if isinstance(param_set, dict) and "names" in param_set:
        return list(param_set["names"])
    n = param_set.n_active() if hasattr(param_set, "n_active") else 2
    return [f"theta_{i}" for i in range(n)]


class InternalIncompressibleForwardModel(ForwardModelBase):
    """Wraps rans_sst_py.ForwardModel; the caller keeps the mesh alive."""

    def __init__(self, cpp_forward_model, param_set):
        self._fm = cpp_forward_model
        self._param_set = param_set

    def evaluate(self, theta) -> EvaluationResult:
        return _wrap(self._fm.evaluate(list(theta)))

    def evaluate_batch(self, thetas: Sequence[Sequence[float]]) -> List[EvaluationResult]:
        return [_wrap(r) for r in self._fm.evaluate_batch([list(t) for t in thetas])]

    def parameter_names(self) -> List[str]:
        return _param_names(self._param_set)


class InternalCompressibleForwardModel(InternalIncompressibleForwardModel):
    """Wraps rans_sst_py.CompressibleForwardModel (pressure-based SIMPLE +
    SST + energy, Ma below about 0.5); same contract as the incompressible
    wrapper (reference python/forward_model_interface.py:130-155).  A
    diverged solve returns non-finite predictions, so converged is False."""


class EnsembleForwardModel(ForwardModelBase):
    """One theta per member model (members may have different meshes); every
    member is resident on the device at once and retires on its own."""

    def __init__(self, cpp_forward_models, param_set):
        from solver_bindings import _rs
        self._rs = _rs()
        self._fms = list(cpp_forward_models)
        self._param_set = param_set

    def evaluate(self, theta) -> EvaluationResult:
        return _wrap(self._fms[0].evaluate(list(theta)))

    def evaluate_members(self, thetas: Sequence[Sequence[float]]) -> List[EvaluationResult]:
        if len(thetas) != len(self._fms):
            raise ValueError("one theta per member model")
        return [_wrap(r) for r in self._rs.evaluate_many(self._fms, [list(t) for t in thetas])]

    def parameter_names(self) -> List[str]:
        return _param_names(self._param_set)
