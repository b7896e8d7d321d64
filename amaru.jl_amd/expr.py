"""Vectorised evaluation of the reference's filter and value expressions.

Amaru writes selections and boundary values as Julia expressions over the
point coordinates and time, e.g. ``:(x==0 && y==0)`` or ``:(3*z)``
(evaluated by src/tools/expr.jl).  Here the same text is parsed once with
Python's ``ast`` (``&&``/``||``/``!`` mapped to ``and``/``or``/``not``)
and evaluated over numpy arrays of coordinates.  Only arithmetic,
comparisons, boolean logic, x/y/z/t and a few math functions are allowed.
"""
import ast
import math

import numpy as np

_FUNCS = {"sqrt": np.sqrt, "abs": np.abs, "sin": np.sin, "cos": np.cos, "exp": np.exp, "log": np.log}
_CONST = {"pi": math.pi}
_CMP = {ast.Eq: np.equal, ast.NotEq: np.not_equal, ast.Lt: np.less, ast.LtE: np.less_equal,
        ast.Gt: np.greater, ast.GtE: np.greater_equal}
_BIN = {ast.Add: np.add, ast.Sub: np.subtract, ast.Mult: np.multiply, ast.Div: np.true_divide,
        ast.Pow: np.power, ast.Mod: np.mod}


def _to_python(src):
    s = src.replace("&&", " and ").replace("||", " or ")
    s = s.replace("!=", "\x00").replace("!", " not ").replace("\x00", "!=").replace("^", "**")
    return s


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
            if isinstance(node.op, ast.Not):
                return np.logical_not(v)
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
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _FUNCS:
            return _FUNCS[node.func.id](*[ev(a, env) for a in node.args])
        raise ValueError(f"unsupported expression element {ast.dump(node)} in {src!r}")

    def f(x, y, z, t=0.0):
        x = np.asarray(x, dtype=float)
        env = {"x": x, "y": np.asarray(y, dtype=float), "z": np.asarray(z, dtype=float), "t": t}
        return np.broadcast_to(np.asarray(ev(tree, env)), x.shape)

    return f
