"""The learned part: an MLP surrogate trained on the GPU by the library's own kernels
screens differential-evolution mutants; this shows what it gains over the method
without it (all figures are measured, nothing is assumed of the surrogate's quality).

  * ranking: on held-out random thickness vectors the surrogate's predicted dark-hole
    contrast correlates with the exact one, and its top 5 % are darker on average;
  * search: at the same number of exact candidate propagations, DE whose 512-mutant
    generations are screened down to 32 by the surrogate reaches a darker dark hole
    than plain DE propagating every one of its 64 mutants.
"""
import pytest

from csim_amd import study
from csim_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_surrogate_screening_gain(ctx):
    w = W.workload3(n=512, n_lambda=9)
    sur, tr = study.train_surrogate(ctx, w, n_train=22528, epochs=20)
    print(tr)
    assert tr["label_consistency_max_rel"] < 1e-9          # labels rebuild the exact objective
    assert tr["train_loss"][1] < 0.1 * tr["train_loss"][0]
    assert tr["holdout_spearman"] > 0.15
    assert tr["holdout_top5pct_mean_over_mean"] < 0.98
    res = study.compare_de(ctx, w, sur, eval_budget=24000)
    print(res)
    assert res["screened"]["best"] < res["plain"]["best"]
