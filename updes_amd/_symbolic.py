"""Symbolic values used to lower user operators into kernel descriptors.

The reference calls a user ``diff_operator(x, center, rbf, monomial, fields)``
once per (node, centre) pair under JAX tracing (``updes/assembly.py:116-135``)
and a ``rhs_operator(x, centers, rbf, fields)`` once per internal node
(``assembly.py:467-470``).  Here each operator is called ONCE, vectorised over
the collocation rows, with symbolic stand-ins:

* ``center`` / ``monomial`` are sentinels; ``nodal_value``, ``nodal_gradient``,
  ``nodal_laplacian`` and ``nodal_div_grad`` then return ``OpTerm`` /
  ``OpVec`` objects -- linear combinations of the operator basis
  (value, d/dx, d/dy, Laplacian, d2/dx2, d2/dy2, d2/dxdy) whose coefficients
  are numbers or per-row tensors;
* ``fields`` of an rhs operator is a ``CoeffFields`` proxy: ``fields[:, k]``
  is a reference to the k-th field's coefficients, and ``value`` /
  ``gradient`` / ``laplacian`` of it return ``EvalExpr`` objects.

Anything that is not linear in the basis (a product of two basis terms, a
non-zero constant added to an operator, a transcendental of a field) raises
``NotLowerable``.
"""
import numbers

import numpy as np
import torch

VAL, DX, DY, LAP, DXX, DYY, DXY = range(7)
NOPS = 7


class NotLowerable(TypeError):
    """The operator is not a linear combination of the built-in nodal operators."""


class _Sentinel:
    def __init__(self, name):
        self._name = name

    def __repr__(self):
        return self._name

    def __eq__(self, other):
        return self is other

    def __ne__(self, other):
        return self is not other

    def __hash__(self):
        return id(self)

    def __bool__(self):
        return True


class _MonomialSentinel(_Sentinel):
    def __call__(self, x):
        return OpTerm({VAL: 1.0})


CENTER = _Sentinel("CENTER")
MONOMIAL = _MonomialSentinel("MONOMIAL")


def _is_coef(v):
    if isinstance(v, (OpTerm, OpVec, EvalExpr, EvalVec)):
        return False
    return isinstance(v, numbers.Number) or torch.is_tensor(v) or isinstance(v, np.ndarray) or \
        isinstance(v, np.generic)


def _coef(v):
    if isinstance(v, (np.ndarray, np.generic)):
        return torch.as_tensor(np.asarray(v, dtype=np.float64))
    return v


def _is_zero(v):
    if isinstance(v, numbers.Number):
        return v == 0
    if torch.is_tensor(v):
        return v.numel() == 1 and float(v) == 0.0 if not v.requires_grad else False
    if isinstance(v, (np.ndarray, np.generic)):
        return np.size(v) == 1 and float(v) == 0.0
    return False


def _vec2(v):
    """Split a length-2 leading axis (tuple, list, array, tensor) into components."""
    if isinstance(v, (OpVec, EvalVec)):
        return v[0], v[1]
    if isinstance(v, (list, tuple)) and len(v) == 2:
        return _coef(v[0]), _coef(v[1])
    v = _coef(v)
    if torch.is_tensor(v) and v.shape[0] == 2:
        return v[0], v[1]
    raise NotLowerable("expected a 2-vector, got %r" % (type(v),))


class OpTerm:
    """sum_t c_t op_t  (a linear nodal operator applied to the basis function)."""
    __array_priority__ = 1000

    def __init__(self, c):
        self.c = dict(c)

    # arithmetic ---------------------------------------------------------
    def __add__(self, o):
        if isinstance(o, OpTerm):
            d = dict(self.c)
            for k, v in o.c.items():
                d[k] = d[k] + v if k in d else v
            return OpTerm(d)
        if _is_coef(o) and _is_zero(o):
            return self
        raise NotLowerable("adding a constant or non-linear term to a nodal operator")

    __radd__ = __add__

    def __neg__(self):
        return OpTerm({k: -v for k, v in self.c.items()})

    def __sub__(self, o):
        return self + (-o if isinstance(o, OpTerm) else (-_coef(o) if _is_coef(o) else o.__neg__()))

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if _is_coef(o):
            o = _coef(o)
            return OpTerm({k: v * o for k, v in self.c.items()})
        raise NotLowerable("product of two nodal operator terms")

    __rmul__ = __mul__

    def __truediv__(self, o):
        if _is_coef(o):
            o = _coef(o)
            return OpTerm({k: v / o for k, v in self.c.items()})
        raise NotLowerable("division by a nodal operator term")

    def __rtruediv__(self, o):
        raise NotLowerable("division by a nodal operator term")

    def __pow__(self, p):
        raise NotLowerable("power of a nodal operator term")

    def __getitem__(self, idx):
        # the reference's scalar operators are 0-d; allow x[...] / x[()]
        if idx is Ellipsis or idx == ():
            return self
        raise NotLowerable("indexing a scalar nodal operator term")

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        return _dispatch(func, args, kwargs or {})

    def __array_function__(self, func, types, args, kwargs):
        return _dispatch(func, args, kwargs)

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if method == "__call__":
            return _dispatch(ufunc, inputs, kwargs)
        return NotImplemented

    def __repr__(self):
        return "OpTerm(%r)" % {k: (v if not torch.is_tensor(v) else tuple(v.shape)) for k, v in self.c.items()}


class OpVec:
    """Gradient-like pair (d/dx term, d/dy term)."""
    __array_priority__ = 1000

    def __init__(self, gx, gy):
        self.g = (gx, gy)

    def __getitem__(self, i):
        if isinstance(i, slice) or i is Ellipsis:
            return self
        return self.g[i]

    def __len__(self):
        return 2

    def __iter__(self):
        return iter(self.g)

    @property
    def shape(self):
        return (2,)

    def dot(self, v):
        a, b = _vec2(v)
        return self.g[0] * a + self.g[1] * b

    def __rmatmul__(self, v):
        return self.dot(v)

    def __matmul__(self, v):
        return self.dot(v)

    def __mul__(self, o):
        if isinstance(o, (list, tuple)) or (torch.is_tensor(o) and o.dim() >= 1 and o.shape[0] == 2) or \
                (isinstance(o, np.ndarray) and o.ndim >= 1 and o.shape[0] == 2):
            a, b = _vec2(o)
            return OpVec(self.g[0] * a, self.g[1] * b)
        return OpVec(self.g[0] * o, self.g[1] * o)

    __rmul__ = __mul__

    def __truediv__(self, o):
        return OpVec(self.g[0] / o, self.g[1] / o)

    def __add__(self, o):
        if isinstance(o, OpVec):
            return OpVec(self.g[0] + o.g[0], self.g[1] + o.g[1])
        if _is_coef(o) and _is_zero(o):
            return self
        raise NotLowerable("adding a constant to a gradient term")

    __radd__ = __add__

    def __neg__(self):
        return OpVec(-self.g[0], -self.g[1])

    def __sub__(self, o):
        return self + (-o)

    def sum(self, *a, **k):
        return self.g[0] + self.g[1]

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        return _dispatch(func, args, kwargs or {})

    def __array_function__(self, func, types, args, kwargs):
        return _dispatch(func, args, kwargs)

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if method == "__call__":
            return _dispatch(ufunc, inputs, kwargs)
        return NotImplemented


# ---- field references (rhs operators) --------------------------------------
class CoeffRef:
    """Reference to the coefficients of the k-th rhs field."""

    def __init__(self, k):
        self.k = k


class CoeffFields:
    """Stand-in for the (n, n_args) coefficient matrix passed to rhs operators."""

    def __init__(self, nargs):
        self.nargs = nargs

    @property
    def shape(self):
        return (None, self.nargs)

    def __getitem__(self, idx):
        if isinstance(idx, tuple) and len(idx) == 2:
            rows, col = idx
            if not (rows is Ellipsis or (isinstance(rows, slice) and rows == slice(None))):
                raise NotLowerable("rhs fields may only be indexed as fields[:, k]")
            if isinstance(col, slice):
                return [CoeffRef(k) for k in range(self.nargs)[col]]
            return CoeffRef(int(col) % self.nargs)
        raise NotLowerable("rhs fields may only be indexed as fields[:, k] (got %r)" % (idx,))


class EvalExpr:
    """const + sum over (op, field k) of coeff * (op applied to field k)."""
    __array_priority__ = 1000

    def __init__(self, terms=None, const=0.0):
        self.terms = dict(terms or {})
        self.const = const

    def __add__(self, o):
        if isinstance(o, EvalExpr):
            d = dict(self.terms)
            for k, v in o.terms.items():
                d[k] = d[k] + v if k in d else v
            return EvalExpr(d, self.const + o.const)
        if _is_coef(o):
            return EvalExpr(self.terms, self.const + _coef(o))
        raise NotLowerable("unsupported operand %r in an rhs expression" % (type(o),))

    __radd__ = __add__

    def __neg__(self):
        return EvalExpr({k: -v for k, v in self.terms.items()}, -self.const)

    def __sub__(self, o):
        return self + (-o if isinstance(o, EvalExpr) else -_coef(o))

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if isinstance(o, EvalExpr):
            if not o.terms:
                return self * o.const
            if not self.terms:
                return o * self.const
            raise NotLowerable("product of two field evaluations (non-linear rhs)")
        if _is_coef(o):
            o = _coef(o)
            return EvalExpr({k: v * o for k, v in self.terms.items()}, self.const * o)
        raise NotLowerable("unsupported operand %r in an rhs expression" % (type(o),))

    __rmul__ = __mul__

    def __truediv__(self, o):
        if _is_coef(o):
            o = _coef(o)
            return EvalExpr({k: v / o for k, v in self.terms.items()}, self.const / o)
        raise NotLowerable("division by a field evaluation")

    def __rtruediv__(self, o):
        raise NotLowerable("division by a field evaluation")

    def __pow__(self, p):
        if not self.terms:
            return EvalExpr({}, self.const ** p)
        raise NotLowerable("power of a field evaluation")

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        return _dispatch(func, args, kwargs or {})

    def __array_function__(self, func, types, args, kwargs):
        return _dispatch(func, args, kwargs)

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if method == "__call__":
            return _dispatch(ufunc, inputs, kwargs)
        return NotImplemented


class EvalVec:
    """Gradient of a field: (d/dx expr, d/dy expr)."""
    __array_priority__ = 1000

    def __init__(self, gx, gy):
        self.g = (gx, gy)

    def __getitem__(self, i):
        if isinstance(i, tuple) and len(i) == 2 and i[0] is Ellipsis:
            i = i[1]
        return self.g[i]

    def __len__(self):
        return 2

    def dot(self, v):
        a, b = _vec2(v)
        return self.g[0] * a + self.g[1] * b

    __matmul__ = dot

    def __rmatmul__(self, v):
        return self.dot(v)

    def __mul__(self, o):
        return EvalVec(self.g[0] * o, self.g[1] * o)

    __rmul__ = __mul__

    def __truediv__(self, o):
        return EvalVec(self.g[0] / o, self.g[1] / o)

    def __add__(self, o):
        if isinstance(o, EvalVec):
            return EvalVec(self.g[0] + o.g[0], self.g[1] + o.g[1])
        raise NotLowerable("adding a non-vector to a field gradient")

    __radd__ = __add__

    def __neg__(self):
        return EvalVec(-self.g[0], -self.g[1])

    def __sub__(self, o):
        return self + (-o)

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        return _dispatch(func, args, kwargs or {})

    def __array_function__(self, func, types, args, kwargs):
        return _dispatch(func, args, kwargs)


_SYM = (OpTerm, OpVec, EvalExpr, EvalVec)


def _dispatch(func, args, kwargs):
    """numpy / torch functions applied to symbolic values."""
    name = getattr(func, "__name__", str(func)).strip("_")
    if name in ("dot", "matmul", "inner", "vdot", "rmatmul"):
        a, b = args[0], args[1]
        if isinstance(b, (OpVec, EvalVec)):
            return b.dot(a)
        if isinstance(a, (OpVec, EvalVec)):
            return a.dot(b)
    binops = {"multiply": "mul", "mul": "mul", "rmul": "mul", "add": "add", "radd": "add", "subtract": "sub",
              "sub": "sub", "rsub": "rsub", "divide": "truediv", "true_divide": "truediv", "div": "truediv",
              "truediv": "truediv", "rtruediv": "rtruediv", "rdiv": "rtruediv"}
    if name in binops:
        op = binops[name]
        a, b = args[0], args[1]
        if op == "rsub":                # torch: rsub(a, b) = b - a
            a, b, op = b, a, "sub"
        if op == "rtruediv":
            a, b, op = b, a, "truediv"
        if isinstance(a, _SYM):
            return getattr(a, "__%s__" % op)(b)
        return getattr(b, "__r%s__" % op)(a)
    if name in ("negative", "neg"):
        return -args[0]
    if name == "sum" and isinstance(args[0], (OpVec, EvalVec)):
        return args[0].g[0] + args[0].g[1]
    if name == "sum" and isinstance(args[0], (OpTerm, EvalExpr)):
        return args[0]
    if name in ("nan_to_num", "asarray", "array") and isinstance(args[0], _SYM):
        return args[0]
    if name == "trace":
        raise NotLowerable("trace of a symbolic Hessian is not supported; use nodal_laplacian")
    raise NotLowerable("function %s is not supported on symbolic operator terms" % name)
