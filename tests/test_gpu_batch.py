"""Batched continuation: instances run concurrently on worker streams give the
same results as the same instances run one after another."""
import numpy
import pytest

pytestmark = pytest.mark.gpu


def test_batch_matches_single_runs(hip):
    from transiflow_amd.batch import run_batch, run_instance
    instances = [{'n': 16, 'Re': 350.0}, {'n': 32, 'Re': 120.0}, {'n': 24, 'Re': 600.0},
                 {'n': 40, 'Re': 200.0}, {'n': 16, 'Re': 480.0}, {'n': 32, 'Re': 300.0}]
    single = [run_instance(inst) for inst in instances]
    batched, wall = run_batch(instances, workers=3)
    for inst, s, b in zip(instances, single, batched):
        err = numpy.linalg.norm(b['x'] - s['x']) / numpy.linalg.norm(s['x'])
        print('%s: steps %d/%d, mu %.6f/%.6f, rel diff %.2e' % (inst, b['accepted'], s['accepted'], b['mu'], s['mu'], err))
        assert b['accepted'] == s['accepted']
        assert b['mu'] == pytest.approx(s['mu'], rel=1e-12, abs=1e-9)
        assert err <= 1e-10
        assert abs(b['mu'] - inst['Re']) <= 1e-3 * inst['Re']   # reached the target
