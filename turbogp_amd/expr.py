"""Expression trees of WHERE terms and their compilation to tgp_scan_expr programs.

Mirrors the interpreter's `Expr2` (query_interpreter/types.rs) for the variants
the per-row fallback of build_mask evaluates on numbers (expr.rs:1626-1800):
columns, Int / Float / Date / NULL literals, + - * / %, comparisons, AND / OR /
NOT, unary minus, BETWEEN, IN lists and CASE WHEN.  A tree compiles to the
postfix program the device runs (children first, then the operator); CASE with
several WHENs nests as IF(cond1, then1, IF(cond2, then2, else)).
"""
from dataclasses import dataclass, field
import struct
from typing import Callable, List, Optional, Sequence, Tuple

from . import _capi as A

INT, FLOAT, DATE = A.EX_INT, A.EX_FLOAT, A.EX_DATE  # column kinds (Value2 variants)


@dataclass(frozen=True)
class Col:
    name: str


@dataclass(frozen=True)
class Lit:
    """An Int (Python int), Float (Python float), Date (kind DATE, days) or NULL
    (value None)."""
    value: object
    kind: Optional[int] = None  # None: INT for int, FLOAT for float


@dataclass(frozen=True)
class BinOp:
    op: str  # + - * / % = <> != < <= > >= AND OR
    left: object
    right: object


@dataclass(frozen=True)
class Not:
    expr: object


@dataclass(frozen=True)
class Neg:
    expr: object


@dataclass(frozen=True)
class Between:
    expr: object
    low: object
    high: object
    negated: bool = False


@dataclass(frozen=True)
class InList:
    expr: object
    items: Tuple[object, ...]
    negated: bool = False


@dataclass(frozen=True)
class Case:
    whens: Tuple[Tuple[object, object], ...]
    else_: Optional[object] = None


_BIN = {"+": A.EX_ADD, "-": A.EX_SUB, "*": A.EX_MUL, "/": A.EX_DIV, "%": A.EX_MOD,
        "=": A.EX_EQ, "<>": A.EX_NE, "!=": A.EX_NE, "<": A.EX_LT, "<=": A.EX_LE,
        ">": A.EX_GT, ">=": A.EX_GE, "AND": A.EX_AND, "OR": A.EX_OR}


def columns(e) -> List[str]:
    """The column names an expression reads, in first-use order."""
    out = []

    def walk(x):
        if isinstance(x, Col):
            if x.name not in out:
                out.append(x.name)
        elif isinstance(x, BinOp):
            walk(x.left)
            walk(x.right)
        elif isinstance(x, (Not, Neg)):
            walk(x.expr)
        elif isinstance(x, Between):
            walk(x.expr), walk(x.low), walk(x.high)
        elif isinstance(x, InList):
            walk(x.expr)
            for i in x.items:
                walk(i)
        elif isinstance(x, Case):
            for c, r in x.whens:
                walk(c), walk(r)
            if x.else_ is not None:
                walk(x.else_)
    walk(e)
    return out


def compile_expr(e, col_ref: Callable[[str], Tuple[int, int]]) -> List[Tuple[int, int, int]]:
    """Postfix program [(code, arg, imm)] of e; col_ref(name) -> (column index, kind)."""
    prog = []

    def lit(v, kind):
        if v is None:
            prog.append((A.EX_NULL, 0, 0))
        elif kind == DATE:
            prog.append((A.EX_DATE, 0, int(v)))
        elif kind == FLOAT or (kind is None and isinstance(v, float)):
            prog.append((A.EX_FLOAT, 0, struct.unpack("<q", struct.pack("<d", float(v)))[0]))
        else:
            prog.append((A.EX_INT, 0, int(v)))

    def emit(x):
        if isinstance(x, Col):
            idx, kind = col_ref(x.name)
            prog.append((A.EX_COL, idx, kind))
        elif isinstance(x, Lit):
            lit(x.value, x.kind)
        elif isinstance(x, BinOp):
            emit(x.left)
            emit(x.right)
            prog.append((_BIN[x.op.upper()], 0, 0))
        elif isinstance(x, Not):
            emit(x.expr)
            prog.append((A.EX_NOT, 0, 0))
        elif isinstance(x, Neg):
            emit(x.expr)
            prog.append((A.EX_NEG, 0, 0))
        elif isinstance(x, Between):
            emit(x.expr), emit(x.low), emit(x.high)
            prog.append((A.EX_NOT_BETWEEN if x.negated else A.EX_BETWEEN, 0, 0))
        elif isinstance(x, InList):
            emit(x.expr)
            for i in x.items:
                emit(i)
            prog.append((A.EX_NOT_IN if x.negated else A.EX_IN, len(x.items), 0))
        elif isinstance(x, Case):
            def chain(k):
                if k == len(x.whens):
                    emit(x.else_) if x.else_ is not None else lit(None, None)
                    return
                emit(x.whens[k][0])
                emit(x.whens[k][1])
                chain(k + 1)
                prog.append((A.EX_IF, 0, 0))
            chain(0)
        elif isinstance(x, (int, float)) or x is None:
            lit(x, None)
        else:
            raise TypeError(f"not an expression: {x!r}")
    emit(e)
    return prog
