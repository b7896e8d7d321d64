"""Vectorised evaluation of the reference's filter and value expressions.

Amaru writes selections and boundary values as Julia expressions over the
point coordinates and time, e.g. ``:(x==0 && y==0)`` or ``:(3*z)``,
evaluated by ``evaluate`` (src/tools/expr.jl:130-187) with the operations
of its ``op_dict`` (src/tools/expr.jl:98-125).  Here the same text is
parsed once with Python's ``ast`` (``&&``/``||`` mapped to ``and``/``or``,
``^`` to ``**``) and evaluated over numpy arrays of coordinates with the
same operation set:
- arithmetic ``+ - * / ^`` and ``div`` (truncated quotient);
- comparisons with the reference's absolute tolerance ``arith_tol = 1e-6``:
  ``a == b`` is ``|a - b| < tol``, ``a != b`` is ``|a - b| >= tol``,
  ``a > b`` is ``a > b + tol``, ``a < b`` is ``a < b - tol``,
  ``a >= b`` is ``a > b - tol``, ``a <= b`` is ``a < b + tol``;
- ``&&`` and ``||``;
- ``abs sin cos tan exp log log10 sqrt max min`` and the constant ``pi``.
Anything else (``%``, a unary ``!``, other functions) is refused, as the
reference refuses an operation missing from ``op_dict``.
"""
import ast
import math

import numpy as np

ARITH_TOL = 1e-6  # src/tools/expr.jl:96


def _div(a, b):
    """Julia's div on floats: the quotient truncated toward zero, from the
    exact remainder (div(a, b) = (a - rem(a, b)) / b)."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return np.trunc((a - np.fmod(a, b)) / b)


_FUNCS = {"sqrt": np.sqrt, "abs": np.abs, "sin": np.sin, "cos": np.cos, "tan": np.tan, "exp": np.exp,
          "log": np.log, "log10": np.log10, "max": np.maximum, "min": np.minimum, "div": _div}
_NARGS = {"max": 2, "min": 2, "div": 2}
_CONST = {"pi": math.pi}
_CMP = {ast.Eq: lambda a, b: np.abs(np.subtract(a, b)) < ARITH_TOL,
        ast.NotEq: lambda a, b: np.abs(np.subtract(a, b)) >= ARITH_TOL,
        ast.Lt: lambda a, b: np.less(a, np.subtract(b, ARITH_TOL)),
        ast.LtE: lambda a, b: np.less(a, np.add(b, ARITH_TOL)),
        ast.Gt: lambda a, b: np.greater(a, np.add(b, ARITH_TOL)),
        ast.GtE: lambda a, b: np.greater(a, np.subtract(b, ARITH_TOL))}
_BIN = {ast.Add: np.add, ast.Sub: np.subtract, ast.Mult: np.multiply, ast.Div: np.true_divide,
        ast.Pow: np.power}


def _to_python(src):
    s = src.replace("&&", " and ").replace("||", " or ")
    s = s.replace("!=", "\x00")
    if "!" in s:
        raise ValueError(f"operation ! not allowed in {src!r} (src/tools/expr.jl op_dict)")
    return s.replace("\x00", "!=").replace("^", "**")


def compile_expr(src):
    """Parse an expression string once; returns an evaluator f(x, y, z, t)."""
    if isinstance(src, (int, float, np.floating, np.integer)):
        v = float(src)
        return lambda x, y, z, t=0.0: np.full(np.shape(x), v)
    if callable(src):
        return src
    tree = ast.parse(_to_python(str(src)), mode="eval").body

    def ev(node, env):
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float, bool)):
            return node.value
        if isinstance(node, ast.Name):
            if node.id in env:
                return env[node.id]
            if node.id in _CONST:
                return _CONST[node.id]
            raise ValueError(f"unknown name {node.id!r} in {src!r}")
        if isinstance(node, ast.BoolOp):
            vals = [np.asarray(ev(v, env), dtype=bool) for v in node.values]
            f = np.logical_and if isinstance(node.op, ast.And) else np.logical_or
            out = vals[0]
            for v in vals[1:]:
                out = f(out, v)
            return out
        if isinstance(node, ast.UnaryOp):
            v = ev(node.operand, env)
            if isinstance(node.op, ast.USub):
                return np.negative(v)
            if isinstance(node.op, ast.UAdd):
                return v
        if isinstance(node, ast.BinOp) and type(node.op) in _BIN:
            return _BIN[type(node.op)](ev(node.left, env), ev(node.right, env))
        if isinstance(node, ast.Compare):
            left = ev(node.left, env)
            out = None
            for op, comp in zip(node.ops, node.comparators):
                right = ev(comp, env)
                r = _CMP[type(op)](left, right)
                out = r if out is None else np.logical_and(out, r)
                left = right
            return out
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _FUNCS \
                and len(node.args) == _NARGS.get(node.func.id, 1) and not node.keywords:
            return _FUNCS[node.func.id](*[ev(a, env) for a in node.args])
        raise ValueError(f"unsupported expression element {ast.dump(node)} in {src!r}")

    def f(x, y, z, t=0.0):
        x = np.asarray(x, dtype=float)
        env = {"x": x, "y": np.asarray(y, dtype=float), "z": np.asarray(z, dtype=float), "t": t}
        return np.broadcast_to(np.asarray(ev(tree, env)), x.shape)

    return f


# postfix programs for the device (the operation codes of AMR_OP_* in
# include/amaru_hip.h)
OP = {"const": 0, "x": 1, "y": 2, "z": 3, "t": 4, "neg": 5, "sqrt": 6, "abs": 7, "sin": 8, "cos": 9, "exp": 10,
      "log": 11, "tan": 12, "log10": 13, ast.Add: 14, ast.Sub: 15, ast.Mult: 16, ast.Div: 17, ast.Pow: 18,
      "max": 19, "min": 20, "div": 21}


def compile_program(src):
    """A numeric value expression as a postfix program: (ops (n x 2 int32:
    operation, argument), consts (float64)).  Comparisons and boolean logic
    (selection filters) are not load values and are refused."""
    if callable(src):
        raise ValueError("a Python callable cannot run on the device; give the value as an expression")
    ops, consts = [], []

    def const(v):
        consts.append(float(v))
        ops.append((OP["const"], len(consts) - 1))

    if isinstance(src, (int, float, np.floating, np.integer)):
        const(src)
    else:
        tree = ast.parse(_to_python(str(src)), mode="eval").body

        def emit(node):
            if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) \
                    and not isinstance(node.value, bool):
                const(node.value)
            elif isinstance(node, ast.Name) and node.id in ("x", "y", "z", "t"):
                ops.append((OP[node.id], 0))
            elif isinstance(node, ast.Name) and node.id in _CONST:
                const(_CONST[node.id])
            elif isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
                emit(node.operand)
                if isinstance(node.op, ast.USub):
                    ops.append((OP["neg"], 0))
            elif isinstance(node, ast.BinOp) and type(node.op) in OP:
                emit(node.left)
                emit(node.right)
                ops.append((OP[type(node.op)], 0))
            elif isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _FUNCS \
                    and len(node.args) == _NARGS.get(node.func.id, 1) and not node.keywords:
                for a in node.args:
                    emit(a)
                ops.append((OP[node.func.id], 0))
            else:
                raise ValueError(f"not a numeric load expression: {ast.dump(node)} in {src!r}")
        emit(tree)
    return np.ascontiguousarray(np.array(ops, dtype=np.int32).reshape(-1, 2)), np.array(consts, dtype=float)
