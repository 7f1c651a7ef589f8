"""Checkpoint and history formats of the reference, without unpickling.

- `checkpoint_save(state, path)` / `checkpoint_load(state, path, ...)`:
  the reference's state checkpoint (`util_op.py:976-1014`), a dict
  {'fields': {name: array}, 'weights': {name: [array, ...]}}, stored as a
  pickle (the reference's format) or, for a path ending in `.npz`, as an
  npz archive with keys 'fields/<name>' and 'weights/<name>/<i>'.
- `read_pickle(path_or_bytes)`: reads such pickles, including the ones the
  reference writes, by interpreting the opcode stream itself. It builds
  only dicts, lists, tuples, numbers, strings, bytes and NumPy arrays,
  dtypes and scalars; any other global in the stream is rejected with
  ValueError. Nothing named by the file is imported or called.
- `History`: the reference's `util_history.History` (`util_history.py`):
  columns appended per entry, `commit()` checks that every column has a
  value, `write()` emits `train.csv` (header on the first write, values with
  `str()`, comma-separated), `save()` / `load()` keep the columns.

Device tensors in a state are copied to host arrays on save; loaded arrays
are host NumPy arrays (callers move them to the device).
"""

import pickle
import pickletools

import numpy as np

__all__ = ['checkpoint_save', 'checkpoint_load', 'read_pickle', 'History']


def _host(a):
    if hasattr(a, 'detach'):
        a = a.detach().cpu().numpy()
    return np.array(a)


def checkpoint_save(state, path):
    """Write state.fields and state.weights (util_op.py:976-990)."""
    fields = {k: _host(v) for k, v in state.fields.items()}
    weights = {k: [_host(w) for w in v]
               for k, v in getattr(state, 'weights', {}).items()}
    if str(path).endswith('.npz'):
        arrs = {'fields/' + k: v for k, v in fields.items()}
        for k, ws in weights.items():
            for i, w in enumerate(ws):
                arrs['weights/{}/{}'.format(k, i)] = w
        np.savez(path, **arrs)
        return
    with open(path, 'wb') as f:
        pickle.dump({'fields': fields, 'weights': weights}, f)


def _read_checkpoint(path):
    if str(path).endswith('.npz'):
        fields, weights = {}, {}
        with np.load(path, allow_pickle=False) as z:
            for key in z.files:
                parts = key.split('/')
                if parts[0] == 'fields' and len(parts) == 2:
                    fields[parts[1]] = z[key]
                elif parts[0] == 'weights' and len(parts) == 3:
                    weights.setdefault(parts[1], {})[int(parts[2])] = z[key]
        weights = {k: [v[i] for i in sorted(v)] for k, v in weights.items()}
        return {'fields': fields, 'weights': weights}
    return read_pickle(path)


def checkpoint_load(state, path, fields_to_load=None, weights_to_load=None,
                    skip_missing=True):
    """Assign checkpointed fields and weights to `state` (util_op.py:
    993-1014): all of them by default, or the listed names; a listed name
    missing from the file raises RuntimeError unless skip_missing."""
    s = _read_checkpoint(path)
    fields = s.get('fields', None)
    if fields is not None:
        for k in (fields.keys() if fields_to_load is None
                  else fields_to_load):
            if k not in fields:
                if not skip_missing:
                    raise RuntimeError('Missing field {}'.format(k))
                continue
            state.fields[k] = fields[k]
    weights = s.get('weights', None)
    if weights is not None:
        if not hasattr(state, 'weights'):
            state.weights = {}
        for k in (weights.keys() if weights_to_load is None
                  else weights_to_load):
            if k not in weights:
                if not skip_missing:
                    raise RuntimeError('Missing weights {}'.format(k))
                continue
            state.weights[k] = weights[k]
    return state


# -- non-executing pickle reader --------------------------------------------

_MARK = object()

# Globals a NumPy array / dtype / scalar pickle names, in NumPy 1.x and 2.x
# module paths, and the codec protocol-2 pickles use for bytes.
_ALLOWED = {
    ('numpy.core.multiarray', '_reconstruct'): 'reconstruct',
    ('numpy._core.multiarray', '_reconstruct'): 'reconstruct',
    ('numpy', 'ndarray'): 'ndarray',
    ('numpy', 'dtype'): 'dtype',
    ('numpy.core.multiarray', 'scalar'): 'scalar',
    ('numpy._core.multiarray', 'scalar'): 'scalar',
    ('numpy.core.numeric', '_frombuffer'): 'frombuffer',
    ('numpy._core.numeric', '_frombuffer'): 'frombuffer',
    ('_codecs', 'encode'): 'encode',
    ('__builtin__', 'bytes'): 'bytes',   # protocol 2 empty bytes: bytes()
    ('builtins', 'bytes'): 'bytes',
}


class _Global:
    def __init__(self, kind):
        self.kind = kind


class _ArrayStub:
    """Result of _reconstruct(ndarray, ...) before its BUILD state."""
    value = None


class _DtypeStub:
    def __init__(self, descr):
        self.descr = descr
        self.value = np.dtype(descr)


def _as_dtype(d):
    if isinstance(d, _DtypeStub):
        return d.value
    if isinstance(d, np.dtype):
        return d
    raise ValueError('pickle: expected a dtype')


def _reduce(fn, args):
    if not isinstance(fn, _Global):
        raise ValueError('pickle: call of a non-allowlisted object')
    k = fn.kind
    if k == 'reconstruct':
        if not (len(args) == 3 and isinstance(args[0], _Global)
                and args[0].kind == 'ndarray'):
            raise ValueError('pickle: unexpected ndarray reconstruction')
        return _ArrayStub()
    if k == 'dtype':
        descr = args[0]
        if not isinstance(descr, str):
            raise ValueError('pickle: unsupported dtype descriptor')
        return _DtypeStub(descr)
    if k == 'scalar':
        dt = _as_dtype(args[0])
        raw = args[1] if len(args) > 1 else b''
        if not isinstance(raw, (bytes, bytearray)):
            raise ValueError('pickle: object scalars are not supported')
        return np.frombuffer(bytes(raw), dt)[0]
    if k == 'frombuffer':
        buf, dt, shape, order = args
        a = np.frombuffer(bytes(buf), _as_dtype(dt))
        return a.reshape(shape, order=order).copy()
    if k == 'encode':
        s, enc = args
        if enc != 'latin1':
            raise ValueError('pickle: unexpected codec {}'.format(enc))
        return s.encode(enc)
    if k == 'bytes':
        if args:
            raise ValueError('pickle: bytes() with arguments')
        return b''
    raise ValueError('pickle: {} cannot be called'.format(k))


def _build(obj, state):
    if isinstance(obj, _DtypeStub):
        endian = state[1]
        if state[3] is not None or state[4] is not None:
            raise ValueError('pickle: structured dtypes are not supported')
        if endian in ('<', '>'):
            obj.value = obj.value.newbyteorder(endian)
        return obj
    if isinstance(obj, _ArrayStub):
        _, shape, dt, fortran, raw = state
        dt = _as_dtype(dt)
        if not isinstance(raw, (bytes, bytearray)):
            raise ValueError('pickle: object arrays are not supported')
        a = np.frombuffer(bytes(raw), dt) if len(raw) else np.zeros(0, dt)
        obj.value = a.reshape(shape, order='F' if fortran else 'C').copy()
        return obj
    raise ValueError('pickle: BUILD on an unsupported object')


def _resolve(x):
    """Replace stubs by their values (recursively through containers)."""
    if isinstance(x, (_ArrayStub, _DtypeStub)):
        return x.value
    if isinstance(x, dict):
        return {_resolve(k): _resolve(v) for k, v in x.items()}
    if isinstance(x, list):
        return [_resolve(v) for v in x]
    if isinstance(x, tuple):
        return tuple(_resolve(v) for v in x)
    return x


def read_pickle(src):
    """Contents of a pickle of containers, numbers, strings and NumPy
    arrays, by interpreting its opcodes (see the module docstring)."""
    data = src if isinstance(src, (bytes, bytearray)) else \
        open(src, 'rb').read()
    stack, memo = [], {}

    def pop_mark():
        i = len(stack) - 1
        while stack[i] is not _MARK:
            i -= 1
        items = stack[i + 1:]
        del stack[i:]
        return items

    for op, arg, _ in pickletools.genops(data):
        n = op.name
        if n in ('PROTO', 'FRAME'):
            continue
        if n == 'STOP':
            return _resolve(stack.pop())
        if n == 'MARK':
            stack.append(_MARK)
        elif n == 'EMPTY_DICT':
            stack.append({})
        elif n == 'EMPTY_LIST':
            stack.append([])
        elif n == 'EMPTY_TUPLE':
            stack.append(())
        elif n in ('BINUNICODE', 'SHORT_BINUNICODE', 'BINUNICODE8',
                   'UNICODE', 'BINSTRING', 'SHORT_BINSTRING', 'STRING'):
            stack.append(arg)
        elif n in ('BINBYTES', 'SHORT_BINBYTES', 'BINBYTES8', 'BYTEARRAY8'):
            stack.append(bytes(arg))
        elif n in ('BININT', 'BININT1', 'BININT2', 'LONG1', 'LONG4', 'INT',
                   'LONG', 'BINFLOAT', 'FLOAT'):
            stack.append(arg)
        elif n == 'NONE':
            stack.append(None)
        elif n == 'NEWTRUE':
            stack.append(True)
        elif n == 'NEWFALSE':
            stack.append(False)
        elif n == 'TUPLE':
            stack.append(tuple(pop_mark()))
        elif n in ('TUPLE1', 'TUPLE2', 'TUPLE3'):
            k = int(n[-1])
            items = stack[-k:]
            del stack[-k:]
            stack.append(tuple(items))
        elif n == 'LIST':
            stack.append(list(pop_mark()))
        elif n == 'DICT':
            items = pop_mark()
            stack.append(dict(zip(items[::2], items[1::2])))
        elif n == 'APPEND':
            v = stack.pop()
            stack[-1].append(v)
        elif n == 'APPENDS':
            items = pop_mark()
            stack[-1].extend(items)
        elif n == 'SETITEM':
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif n == 'SETITEMS':
            items = pop_mark()
            d = stack[-1]
            for k, v in zip(items[::2], items[1::2]):
                d[k] = v
        elif n in ('BINPUT', 'LONG_BINPUT', 'PUT'):
            memo[arg] = stack[-1]
        elif n == 'MEMOIZE':
            memo[len(memo)] = stack[-1]
        elif n in ('BINGET', 'LONG_BINGET', 'GET'):
            stack.append(memo[arg])
        elif n in ('GLOBAL', 'STACK_GLOBAL'):
            if n == 'GLOBAL':
                mod, name = arg.split(' ', 1)
            else:
                name = stack.pop()
                mod = stack.pop()
            kind = _ALLOWED.get((mod, name))
            if kind is None:
                raise ValueError('pickle: global {}.{} is not allowed'
                                 .format(mod, name))
            stack.append(_Global(kind))
        elif n == 'REDUCE':
            args = stack.pop()
            fn = stack.pop()
            stack.append(_reduce(fn, args))
        elif n == 'BUILD':
            state = stack.pop()
            stack[-1] = _build(stack[-1], state)
        else:
            raise ValueError('pickle: opcode {} is not supported'.format(n))
    raise ValueError('pickle: no STOP opcode')


# -- history ------------------------------------------------------------------

class History:
    """Columns of per-epoch values written as CSV (util_history.py)."""

    def __init__(self, csvpath=None, warmup=0):
        self.data = {}
        self.count = 0
        self.warmup = warmup
        self.csvcount = 0
        self.csvpath = csvpath
        self.csvkeys = None
        self.csvfile = open(csvpath, 'w') if csvpath is not None else None

    @staticmethod
    def _none_like(value):
        if value is None:
            return None
        if isinstance(value, (float, np.floating)):
            return 0.
        if isinstance(value, (int, np.integer)):
            return 0
        return None

    def append(self, key, value=None):
        if not (value is None or isinstance(
                value, (int, float, str, np.floating, np.integer))):
            raise TypeError('Unexpected type: ' + str(type(value)))
        if key not in self.data:
            if value is None:
                raise ValueError('first value of a column must be given')
            self.data[key] = [self._none_like(value)] * self.count
        if value is None:
            if not self.data[key]:
                raise ValueError('Expected non-empty column ' + key)
            value = self._none_like(self.data[key][-1])
        self.data[key].append(value)

    def append_dict(self, newdict):
        for k, v in newdict.items():
            self.append(k, v)

    def get(self, key, default=None):
        return self.data.get(key, default)

    def commit(self):
        """Finish the current entry; every column must have a value."""
        maxlen = max(len(v) for v in self.data.values())
        fail = ''.join(k + ',' for k, v in self.data.items()
                       if len(v) < maxlen)
        if fail:
            raise RuntimeError('Missing values for columns: ' + fail)
        self.count += 1

    def write(self, nocommit=False):
        """Commit the entry (unless nocommit) and write pending rows."""
        if not nocommit:
            self.commit()
        if self.count <= self.warmup or self.csvfile is None:
            return
        if self.csvkeys is not None and \
                len(self.data.keys()) != len(self.csvkeys):
            newkeys = list(set(self.data.keys()) - set(self.csvkeys))
            raise RuntimeError('Unexpected keys in history: {}'.format(
                newkeys))
        if self.csvcount == 0:
            self.csvkeys = list(self.data.keys())
            self.csvfile.write(','.join(self.csvkeys) + '\n')
        while self.csvcount < self.count:
            row = [self.data[key][self.csvcount] for key in self.data]
            self.csvfile.write(','.join(map(str, row)) + '\n')
            self.csvcount += 1
        self.csvfile.flush()

    def save(self, path):
        with open(path, 'wb') as f:
            pickle.dump(self.data, f)

    def load(self, path):
        """Replace the columns by those saved at `path` and write them."""
        self.data = read_pickle(path)
        self.csvkeys = list(self.data.keys())
        self.count = len(next(iter(self.data.values())))
        self.write(nocommit=True)

    def close(self):
        if self.csvfile is not None:
            self.csvfile.close()
