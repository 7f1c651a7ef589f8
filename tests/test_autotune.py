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
