"""The trained closure against the plain coarse solver: a network trained on box-
filtered 64^3 runs (one seed) makes a 16^3 run of another seed closer, shell by
shell, to the box-filtered 64^3 run than the same coarse run without it."""
import pytest

pytestmark = pytest.mark.gpu


def test_trained_closure_beats_the_plain_coarse_run():
    from laps_amd.closure import experiment
    r = experiment(n_fine=64, r=4, train_seeds=(31,), test_seed=32, nsteps=192, every=16, amp=0.4, di=0.05,
                   visc=5e-3, train_steps=1500, log=lambda m: print(m, flush=True))
    plain, cl = r["plain"], r["closure"]
    print(f"shell deviations plain {plain['shell_dev']}\n            closure {cl['shell_dev']}")
    assert r["loss_last"] < 0.6 * r["loss_first"]
    assert cl["mean_shell_dev"] < plain["mean_shell_dev"]
