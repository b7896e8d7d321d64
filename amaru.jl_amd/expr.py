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


# postfix programs for the device (the operation codes of AMR_OP_* in
# include/amaru_hip.h)
OP = {"const": 0, "x": 1, "y": 2, "z": 3, "t": 4, "neg": 5, "sqrt": 6, "abs": 7, "sin": 8, "cos": 9, "exp": 10,
      "log": 11, ast.Add: 12, ast.Sub: 13, ast.Mult: 14, ast.Div: 15, ast.Pow: 16, ast.Mod: 17}


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
                    and len(node.args) == 1:
                emit(node.args[0])
                ops.append((OP[node.func.id], 0))
            else:
                raise ValueError(f"not a numeric load expression: {ast.dump(node)} in {src!r}")
        emit(tree)
    return np.ascontiguousarray(np.array(ops, dtype=np.int32).reshape(-1, 2)), np.array(consts, dtype=float)
