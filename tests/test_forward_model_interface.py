"""The Python forward-model protocol (reference python/
forward_model_interface.py:42-155) on stand-in C++ models: the wrappers'
convergence flag, status, metadata and parameter names.  The device-backed
models behind them are exercised in test_gpu_comp.py and test_gpu_parity.py."""
import math
from types import SimpleNamespace

import pytest

import forward_model_interface as FMI


class _StubModel:
    def __init__(self, preds, status="SolverStatus.Converged"):
        self.preds, self.status, self.calls = preds, status, []

    def _res(self, theta):
        self.calls.append(list(theta))
        return SimpleNamespace(predictions=list(self.preds), status=self.status, simple_iters=7, log_lik=-1.5,
                               latency=0.01)

    def evaluate(self, theta):
        return self._res(theta)

    def evaluate_batch(self, thetas):
        return [self._res(t) for t in thetas]


@pytest.mark.parametrize("cls", [FMI.InternalIncompressibleForwardModel, FMI.InternalCompressibleForwardModel])
def test_wrapper_converged_flag_and_status(cls):
    m = cls(_StubModel([0.004, 0.0035]), {"names": ["a1", "betaStar"]})
    r = m.evaluate((0.31, 0.09))
    assert r.converged and bool(r) and r.predictions == [0.004, 0.0035]
    assert r.status == "SolverStatus.Converged"
    assert r.metadata == {"simple_iters": 7, "log_lik": -1.5, "latency_s": 0.01}
    assert m._fm.calls == [[0.31, 0.09]]
    assert m.parameter_names() == ["a1", "betaStar"]
    assert m.precomputed_ensemble() is None
    rs = m.evaluate_batch([(0.3, 0.09), (0.32, 0.088)])
    assert [r.converged for r in rs] == [True, True] and len(m._fm.calls) == 3


@pytest.mark.parametrize("preds", [[], [0.004, math.nan], [math.inf]])
def test_wrapper_nonfinite_or_empty_is_not_converged(preds):
    r = FMI.InternalCompressibleForwardModel(_StubModel(preds, "SolverStatus.Diverged"), {}).evaluate([0.3, 0.09])
    assert not r.converged and not bool(r)
    assert r.status == "SolverStatus.Diverged"


def test_parameter_names_fallbacks():
    class _Set:
        def n_active(self):
            return 3
    assert FMI._param_names({}) == ["theta_0", "theta_1"]
    assert FMI._param_names({"a1": 0.31}) == ["theta_0", "theta_1"]
    assert FMI._param_names(_Set()) == ["theta_0", "theta_1", "theta_2"]
    assert FMI._param_names(SimpleNamespace(active_names=lambda: ["a1"])) == ["a1"]


class _BareResult:
    """A C++ result with only the two attributes the protocol requires."""
    def __init__(self, preds, status):
        self.predictions, self.status = preds, status


class _BareModel:
    def __init__(self, fail=False):
        self.fail = fail

    def evaluate(self, theta):
        return _BareResult([], "diverged") if self.fail else _BareResult([0.01 * t for t in theta], "converged")


@pytest.mark.parametrize("cls", [FMI.InternalIncompressibleForwardModel, FMI.InternalCompressibleForwardModel])
def test_wrapper_needs_only_predictions_and_status(cls):
    ok = cls(_BareModel(), {"names": ["a1", "betaStar"]}).evaluate([0.31, 0.09])
    assert isinstance(ok, FMI.EvaluationResult)
    assert ok.predictions == pytest.approx([0.0031, 0.0009]) and ok.converged and ok.status == "converged"
    assert ok.metadata == {}
    bad = cls(_BareModel(fail=True), {}).evaluate([0.31, 0.09])
    assert bad.predictions == [] and not bad.converged and bad.status == "diverged"


def test_evaluation_result_defaults_and_truth():
    r = FMI.EvaluationResult([1.0, 2.0])
    assert r.status == "ok" and r.converged is True and r.metadata == {} and bool(r)
    assert not FMI.EvaluationResult([], converged=True)
    assert not FMI.EvaluationResult([0.003], converged=False)


def test_base_is_abstract():
    with pytest.raises(TypeError):
        FMI.ForwardModelBase()
