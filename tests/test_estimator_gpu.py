"""The learned estimator end to end on the GPU: trained by the device MLP
kernels on plain-engine labels, it must beat the estimate without the
network on held-out plans and never change a result."""
import pytest

from bespokeolap_artifacts_amd import estimator as E

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def result(engine):
    return E.experiment(engine, n_train_schemas=8, train_plans=300, n_eval_schemas=1,
                        eval_plans=120, eval_fact_rows=4_000_000, seed=3, epochs=40, reps=1)


def test_results_are_bit_identical(result):
    assert result["bit_identical"]


def test_training_converges(result):
    assert result["final_loss"] < 0.1 * result["first_loss"]


def test_network_beats_the_estimate_without_it(result):
    qn, qb = result["q_error"], result["q_error_without_network"]
    for k in ("dim_pass", "matches", "groups"):
        assert qn[k]["p90"] < qb[k]["p90"], (k, qn[k], qb[k])
    assert qn["matches"]["p50"] < qb["matches"]["p50"]
    assert result["fallback_rate"] <= max(0.05, result["fallback_rate_without_network"])


def test_learned_sizing_skips_the_counting_passes(result):
    # Hints replace two counting passes and two host round trips per plan.
    assert result["exact_over_learned"] > 1.0


def test_library_estimator_gives_the_same_hints(result):
    """The saved estimator loaded into the library (C featurisation, the same
    MLP kernels, the same policy) sizes every held-out plan as Python does."""
    assert result["library_hints_same"] == 1.0
    assert result["library_estimator_ms_per_plan"] > 0
