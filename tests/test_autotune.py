"""Autotuner choice table (CPU: the stored path times nothing)."""

import json

from odil_amd import autotune


def _fresh(monkeypatch, path):
    monkeypatch.setenv('ODIL_AUTOTUNE_FILE', str(path))
    monkeypatch.setattr(autotune, '_CACHE', {})
    monkeypatch.setattr(autotune, '_FILE', {'path': None, 'table': {}})


def test_stored_choice_is_used_without_timing(tmp_path, monkeypatch):
    path = tmp_path / 'tune.json'
    key = ('apply_sym', (513, 513), 72, 24, 'torch.float32', True)
    path.write_text(json.dumps({repr(key): repr((1, 2))}))
    _fresh(monkeypatch, path)

    def run(c):
        raise AssertionError('a stored key must not be timed')
    cand = [(L, p) for L in (0, 1, 2) for p in (1, 2)]
    assert autotune.pick(key, cand, run) == (1, 2)
    assert autotune.cached(key) == (1, 2)


def test_stale_entry_is_ignored_and_new_choices_written(tmp_path,
                                                        monkeypatch):
    path = tmp_path / 'tune.json'
    key = ('poisson_batch', (17, 33), 'torch.float64')
    path.write_text(json.dumps({repr(key): repr((64, 64))}))
    _fresh(monkeypatch, path)
    # (64, 64) is not among the candidates: the table cannot name it.
    assert autotune._from_file(key, [(8, 32), (16, 16)]) is None
    autotune._to_file(key, (16, 16))
    autotune._to_file(('other', 1), 4)
    table = json.loads(path.read_text())
    assert table == {repr(key): '(16, 16)', repr(('other', 1)): '4'}


def test_in_situ_choice_is_installed_and_overrides(tmp_path, monkeypatch):
    path = tmp_path / 'tune.json'
    key = ('apply_sym', (513, 513), 72, 51, 'torch.float64', True)
    mark = ('in_situ', ) + key
    # The isolated choice (-2, 1) and the in-situ one (1, 1) both stored:
    # the in-situ one is installed without timing and replaces the other.
    path.write_text(json.dumps({repr(key): repr((-2, 1)),
                                repr(mark): repr((1, 1))}))
    _fresh(monkeypatch, path)
    installed = []

    def run():
        raise AssertionError('a stored in-situ key must not be timed')
    cand = [(1, 1), (2, 1), (-2, 1)]
    assert autotune.pick_in_situ(key, cand, installed.append, run) == (1, 1)
    assert installed == [(1, 1)]
    assert autotune.cached(key) == (1, 1)
    assert autotune.pick(key, cand, lambda c: None) == (1, 1)


def test_in_situ_off_with_autotune_off(monkeypatch):
    monkeypatch.setenv('ODIL_AUTOTUNE', '0')
    monkeypatch.setattr(autotune, '_CACHE', {})

    def fail(*a):
        raise AssertionError('tuning is off')
    assert autotune.pick_in_situ(('k', 1), [(1, 1), (2, 1)], fail,
                                 fail) is None
