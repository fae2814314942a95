"""The command-line run (python -m zmpo_dmrg_amd.run) on a small case: a
Normal2 ramp, the environment budget (environments moved to host memory
and back), and the FP64 / FP32 comparison sweep from one checkpoint."""
import json

import pytest

pytestmark = pytest.mark.gpu


def test_run_cli_normal2_budget_and_compare(gpu, capsys):
    from zmpo_dmrg_amd import run
    rc = run.main(['--K', '6', '--D', '16', '--schedule', 'normal', '--sweeps', '3',
                   '--tol', '1e-8', '--budget-gb', '1', '--env-budget-gb', '0.00001',
                   '--compare', '1'])
    assert rc == 0
    out = capsys.readouterr().out
    assert 'environment moves' in out                  # the budget moved environments
    res = json.loads(out.strip().splitlines()[-1])
    assert len(res['energies']) == 3
    cmp_ = res['compare']
    assert set(cmp_) == {'0', '1'}
    assert cmp_['1']['nsigma_low'] > 0 and cmp_['1']['nsigma_f64'] > 0
    # the first step solves the same H_eff in both runs: the FP32 form's
    # FP64 Rayleigh quotient agrees to well below the 1e-6 Ha target
    assert cmp_['1']['dE_first_step'] < 1e-7
    assert cmp_["1"]["dE_sweep"] < 1e-5
