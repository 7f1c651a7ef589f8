"""Checkpoint / history formats (util_op.py:976-1014, util_history.py)."""

import os
import pickle

import numpy as np
import pytest

from odil_amd import checkpoint as ck

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


class _State:
    def __init__(self, fields=None, weights=None):
        self.fields = dict(fields or {})
        self.weights = dict(weights or {})


def _sample(rng):
    return _State({'u': rng.standard_normal((5, 7)),
                   'v': rng.standard_normal((5, 7)).astype(np.float32),
                   'n': np.arange(6, dtype=np.int64).reshape(2, 3)},
                  {'w': [rng.standard_normal(3), rng.standard_normal((2, 2))]})


@pytest.mark.parametrize('suffix', ['.pickle', '.npz'])
def test_checkpoint_round_trip(tmp_path, suffix):
    rng = np.random.default_rng(0)
    s = _sample(rng)
    path = str(tmp_path / ('state' + suffix))
    ck.checkpoint_save(s, path)
    t = ck.checkpoint_load(_State(), path)
    for k in s.fields:
        np.testing.assert_array_equal(t.fields[k], s.fields[k])
        assert t.fields[k].dtype == s.fields[k].dtype
    for a, b in zip(t.weights['w'], s.weights['w']):
        np.testing.assert_array_equal(a, b)
    # selected names, missing names
    t = ck.checkpoint_load(_State(), path, fields_to_load=['u'])
    assert list(t.fields) == ['u']
    ck.checkpoint_load(_State(), path, fields_to_load=['q'])
    with pytest.raises(RuntimeError):
        ck.checkpoint_load(_State(), path, fields_to_load=['q'],
                           skip_missing=False)


@pytest.mark.parametrize('protocol', [2, 3, 4, 5])
def test_read_pickle_matches_pickle_module(protocol):
    """Pickles written by Python's pickler (as the reference writes its
    checkpoints and histories) read back equal without unpickling."""
    rng = np.random.default_rng(protocol)
    obj = {'fields': {'u': rng.standard_normal((4, 3)),
                      'f': np.asfortranarray(rng.standard_normal((3, 5))),
                      'i': np.arange(4, dtype=np.int32),
                      'z': np.zeros((0, 2)),
                      's': np.float64(2.5) * np.ones(())},
           'weights': {'w': [rng.standard_normal(2)]},
           'loss': [np.float64(0.25), 1, 2.0, 'x', None, True],
           'shape': (1, 2, 3), 'big': 2**70}
    obj['again'] = obj['fields']['u']  # memo reuse
    got = ck.read_pickle(pickle.dumps(obj, protocol=protocol))
    for k in obj['fields']:
        np.testing.assert_array_equal(got['fields'][k], obj['fields'][k])
        assert got['fields'][k].dtype == obj['fields'][k].dtype
    np.testing.assert_array_equal(got['again'], obj['fields']['u'])
    assert got['loss'] == obj['loss'] and got['shape'] == obj['shape']
    assert got['big'] == obj['big']
    np.testing.assert_array_equal(got['weights']['w'][0],
                                  obj['weights']['w'][0])


def test_read_pickle_rejects_other_globals():
    with pytest.raises(ValueError):
        ck.read_pickle(pickle.dumps({'f': os.getcwd}))
    with pytest.raises(ValueError):
        ck.read_pickle(pickle.dumps(np.array([object()], dtype=object)))


@pytest.mark.skipif(not os.path.isdir('/root/reference/code/cases'),
                    reason='reference checkout not present')
def test_reference_checkpoint_reads_as_fixture():
    """The reference's own checkpoint (written by its checkpoint_save) reads
    back, without unpickling, as the committed fixture extracted from it."""
    path = ('/root/reference/code/cases/cavity_stream/'
            'cavity_N128_Re100.pickle')
    s = ck.checkpoint_load(_State(), path)
    ref = np.load(os.path.join(GOLD, 'cavity_N128_Re100.npz'))
    for k in 'uvp':
        np.testing.assert_array_equal(s.fields[k], ref[k])


def test_history_csv(tmp_path):
    path = str(tmp_path / 'train.csv')
    h = ck.History(path)
    h.append('epoch', 0)
    h.append('loss', 1.5)
    h.write()
    h.append('epoch', 1)
    h.append('loss', np.float64(0.25))
    h.write()
    h.append('epoch', 2)
    with pytest.raises(RuntimeError):
        h.commit()  # 'loss' has no value for this entry
    h.append('loss')  # None -> zero of the column's type
    h.write()
    h.close()
    assert open(path).read() == 'epoch,loss\n0,1.5\n1,0.25\n2,0.0\n'
    p = str(tmp_path / 'hist.pickle')
    h.save(p)
    h2 = ck.History(str(tmp_path / 'again.csv'))
    h2.load(p)
    h2.close()
    assert open(str(tmp_path / 'again.csv')).read() == open(path).read()
