"""A data-only reader for pickles of numpy data (the reference's gold vectors and Library files).

It walks the pickle opcode stream itself (pickletools.genops parses, it does not execute) and builds plain
Python values. It never resolves a GLOBAL to a Python object: the only globals it accepts are numpy's array,
dtype and scalar reconstructors, which it replaces by its own array builders (np.frombuffer on the pickled
bytes), and a named list of classes, which become `Stub` records holding their pickled state as plain data.
Any other global or opcode raises. Nothing in the file can run code.
"""
import pickletools

import numpy as np

_ARRAY_GLOBALS = {('numpy.core.multiarray', '_reconstruct'), ('numpy._core.multiarray', '_reconstruct')}
_NDARRAY = {('numpy', 'ndarray')}
_DTYPE = {('numpy', 'dtype')}
_SCALAR = {('numpy.core.multiarray', 'scalar'), ('numpy._core.multiarray', 'scalar')}


class Stub:
    """A pickled instance of an allowed class: its class name and its state (plain data)."""

    def __init__(self, cls):
        self.cls = cls
        self.state = None

    def __repr__(self):
        return 'Stub(%s)' % self.cls


class _Global:
    def __init__(self, module, name):
        self.key = (module, name)


class _ArrayBuilder:
    def __init__(self):
        self.value = None

    def setstate(self, state):
        # (version, shape, dtype, is_fortran, raw bytes or list)
        _, shape, dt, fortran, raw = state
        dt = dt.value if isinstance(dt, _DTypeBuilder) else np.dtype(dt)
        if isinstance(raw, (bytes, bytearray)):
            a = np.frombuffer(bytes(raw), dtype=dt).copy()
        else:
            raise ValueError('object arrays are not read')
        self.value = a.reshape(shape, order='F' if fortran else 'C')


class _DTypeBuilder:
    def __init__(self, code):
        self.value = np.dtype(code)

    def setstate(self, state):
        # (version, byte order, subarray, names, fields, elsize, alignment, flags)
        order = state[1]
        if order in ('<', '>', '='):
            self.value = self.value.newbyteorder(order)


def _reduce(func, args, allowed):
    if not isinstance(func, _Global):
        raise ValueError('REDUCE of a non-global')
    k = func.key
    if k in _ARRAY_GLOBALS:
        return _ArrayBuilder()
    if k in _DTYPE:
        return _DTypeBuilder(args[0])
    if k in _SCALAR:
        dt, raw = args
        dt = dt.value if isinstance(dt, _DTypeBuilder) else np.dtype(dt)
        return np.frombuffer(bytes(raw), dtype=dt)[0]
    raise ValueError('global %s.%s is not allowed' % k)


def _finish(v):
    if isinstance(v, _ArrayBuilder):
        return v.value
    if isinstance(v, _DTypeBuilder):
        return v.value
    if isinstance(v, dict):
        return {_finish(k): _finish(x) for k, x in v.items()}
    if isinstance(v, list):
        return [_finish(x) for x in v]
    if isinstance(v, tuple):
        return tuple(_finish(x) for x in v)
    if isinstance(v, Stub):
        v.state = _finish(v.state)
    return v


def load(path, stub_classes=()):
    """Read a pickle of dicts/lists/tuples/strings/numbers/numpy arrays; instances of `stub_classes`
    ((module, name) pairs) become Stub records."""
    allowed = set(tuple(c) for c in stub_classes)
    with open(path, 'rb') as f:
        data = f.read()
    stack, memo, marks = [], {}, []

    def pop_mark():
        m = marks.pop()
        items = stack[m:]
        del stack[m:]
        return items

    for op, arg, _ in pickletools.genops(data):
        n = op.name
        if n in ('PROTO', 'FRAME'):
            continue
        if n == 'STOP':
            break
        if n == 'MARK':
            marks.append(len(stack))
        elif n in ('EMPTY_DICT',):
            stack.append({})
        elif n == 'EMPTY_LIST':
            stack.append([])
        elif n == 'EMPTY_TUPLE':
            stack.append(())
        elif n in ('BINPUT', 'LONG_BINPUT', 'PUT'):
            memo[arg] = stack[-1]
        elif n == 'MEMOIZE':
            memo[len(memo)] = stack[-1]
        elif n in ('BINGET', 'LONG_BINGET', 'GET'):
            stack.append(memo[arg])
        elif n in ('BINUNICODE', 'SHORT_BINUNICODE', 'UNICODE', 'BINUNICODE8'):
            stack.append(arg)
        elif n in ('BININT', 'BININT1', 'BININT2', 'INT', 'LONG1', 'LONG4', 'LONG'):
            stack.append(arg)
        elif n in ('BINFLOAT', 'FLOAT'):
            stack.append(arg)
        elif n in ('SHORT_BINBYTES', 'BINBYTES', 'BINBYTES8'):
            stack.append(arg)
        elif n == 'NONE':
            stack.append(None)
        elif n == 'NEWTRUE':
            stack.append(True)
        elif n == 'NEWFALSE':
            stack.append(False)
        elif n == 'TUPLE':
            stack.append(tuple(pop_mark()))
        elif n == 'TUPLE1':
            stack[-1:] = [(stack[-1],)]
        elif n == 'TUPLE2':
            stack[-2:] = [tuple(stack[-2:])]
        elif n == 'TUPLE3':
            stack[-3:] = [tuple(stack[-3:])]
        elif n == 'LIST':
            stack.append(list(pop_mark()))
        elif n == 'DICT':
            items = pop_mark()
            stack.append({items[i]: items[i + 1] for i in range(0, len(items), 2)})
        elif n == 'SETITEM':
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif n == 'SETITEMS':
            items = pop_mark()
            d = stack[-1]
            for i in range(0, len(items), 2):
                d[items[i]] = items[i + 1]
        elif n == 'APPEND':
            v = stack.pop()
            stack[-1].append(v)
        elif n == 'APPENDS':
            items = pop_mark()
            stack[-1].extend(items)
        elif n == 'GLOBAL':
            module, name = arg.split(' ', 1)
            stack.append(_Global(module, name))
        elif n == 'STACK_GLOBAL':
            name = stack.pop()
            module = stack.pop()
            stack.append(_Global(module, name))
        elif n == 'REDUCE':
            args = stack.pop()
            func = stack.pop()
            stack.append(_reduce(func, args, allowed))
        elif n == 'NEWOBJ':
            args = stack.pop()
            cls = stack.pop()
            if not isinstance(cls, _Global) or cls.key not in allowed:
                raise ValueError('NEWOBJ of %s is not allowed' % (getattr(cls, 'key', cls),))
            stack.append(Stub('%s.%s' % cls.key))
        elif n == 'BUILD':
            state = stack.pop()
            obj = stack[-1]
            if isinstance(obj, (_ArrayBuilder, _DTypeBuilder)):
                obj.setstate(state)
            elif isinstance(obj, Stub):
                obj.state = state
            else:
                raise ValueError('BUILD on %r' % type(obj))
        else:
            raise ValueError('opcode %s is not read' % n)
    if len(stack) != 1:
        raise ValueError('malformed pickle')
    return _finish(stack[0])
