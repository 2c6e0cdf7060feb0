"""Compile pointwise sympy expressions into the library's register-machine bytecode.

An instruction is one 32-bit word (see include/mol.h, "pointwise programs"):
opcode | dst<<8 | a<<12 | b<<16 | bconst<<20 | imm<<21.  Registers are allocated as a
stack (registers >= dst are free while computing into dst): a leaf writes the register at
the current depth, a binary operation combines the two top registers into the lower one.
The device interprets the code as a stack machine on that discipline (the library checks
it).  Operands are ordered by register need (Sethi-Ullman), so an expression tree needs as
few registers as possible; more than NREG is an error.  The result is register 0.
"""
from __future__ import annotations

from typing import Callable, Dict, List, Tuple

import sympy as sp

NREG = 6   # depth of the device interpreter's shift-register stack (MOL_STACK)

OP = dict(
    CONST=0, FIELD=1, SLOT=2, COORD=3, TIME=4, PARAM=5, AUX=6, DINNER=7, MOV=8,
    ADD=16, SUB=17, MUL=18, DIV=19, POW=20, MIN=21, MAX=22, LT=23, LE=24, GT=25, GE=26,
    EQ=27, NE=28, AND=29, OR=30, SELECT=31, FMA=32,
    NEG=48, ABS=49, SQRT=50, EXP=51, LOG=52, SIN=53, COS=54, TAN=55, TANH=56, SINH=57,
    COSH=58, ATAN=59, ASIN=60, ACOS=61, SIGN=62, POWI=63, ERF=64, LOG10=65, CBRT=66,
    INV=67, SQ=68, EXPM1=69, LOG1P=70, FLOOR=71, CEIL=72, HEAVISIDE=73,
)

_UNARY = {
    sp.exp: "EXP", sp.log: "LOG", sp.sin: "SIN", sp.cos: "COS", sp.tan: "TAN",
    sp.tanh: "TANH", sp.sinh: "SINH", sp.cosh: "COSH", sp.atan: "ATAN", sp.asin: "ASIN",
    sp.acos: "ACOS", sp.sign: "SIGN", sp.erf: "ERF", sp.Abs: "ABS", sp.floor: "FLOOR",
    sp.ceiling: "CEIL", sp.Heaviside: "HEAVISIDE",
}

_REL = {sp.StrictLessThan: "LT", sp.LessThan: "LE", sp.StrictGreaterThan: "GT",
        sp.GreaterThan: "GE", sp.Equality: "EQ", sp.Unequality: "NE"}


class CompileError(ValueError):
    pass


def encode(op: str, d: int = 0, a: int = 0, b: int = 0, bconst: bool = False, imm: int = 0) -> int:
    if not (0 <= d < NREG and 0 <= a < NREG and 0 <= b < NREG):
        raise CompileError("register out of range")
    if imm < 0:
        imm &= 0x7FF
    if imm >= 2048:
        raise CompileError("immediate out of range")
    return (OP[op] | (d << 8) | (a << 12) | (b << 16) | (int(bconst) << 20) | (imm << 21)) & 0xFFFFFFFF


def _is_number(e) -> bool:
    return isinstance(e, (int, float)) or (isinstance(e, sp.Basic) and e.is_number and not e.free_symbols
                                          and not isinstance(e, (sp.Piecewise,)) and e.is_real is not False)


class ConstPool:
    def __init__(self):
        self.values: List[float] = []
        self.index: Dict[Tuple[int, float], int] = {}

    def add(self, v: float) -> int:
        v = float(v)
        key = (hash(v.hex()), v)
        if key not in self.index:
            self.index[key] = len(self.values)
            self.values.append(v)
        i = self.index[key]
        if i >= 2048:
            raise CompileError("too many program constants")
        return i


class Compiler:
    """Compiles expressions whose leaves are resolved by ``leaf(sym) -> (opname, imm)``."""

    def __init__(self, leaf: Callable[[sp.Basic], Tuple[str, int]], cpool: ConstPool):
        self.leaf = leaf
        self.cpool = cpool

    # --- register need -----------------------------------------------------------
    def need(self, e) -> int:
        if _is_number(e) or e.is_Symbol or isinstance(e, sp.Function) and not e.args:
            return 1
        parts = self._operands(e)
        if not parts:
            return 1
        needs = sorted((self.need(p) for p in parts if not _is_number(p)), reverse=True)
        best = 1
        for k, n in enumerate(needs):
            best = max(best, n + k)
        return best

    def _operands(self, e):
        if isinstance(e, sp.Piecewise):
            out = []
            for val, cond in e.args:
                out.append(val)
                if cond is not sp.true:
                    out.append(cond)
            return out
        return list(e.args)

    # --- code generation -----------------------------------------------------------
    def compile(self, expr) -> List[int]:
        expr = sp.sympify(expr)
        code: List[int] = []
        self._gen(expr, 0, code)
        return code

    def _const(self, v) -> int:
        return self.cpool.add(float(v))

    def _gen(self, e, d: int, code: List[int]) -> None:
        if d >= NREG:
            raise CompileError("expression needs more than %d registers; split it" % NREG)
        if e is sp.true:
            code.append(encode("CONST", d, imm=self._const(1.0)))
            return
        if e is sp.false:
            code.append(encode("CONST", d, imm=self._const(0.0)))
            return
        if _is_number(e):
            code.append(encode("CONST", d, imm=self._const(e)))
            return
        if e.is_Symbol or (isinstance(e, sp.Function) and e.func not in _UNARY
                           and not isinstance(e, (sp.Min, sp.Max, sp.Piecewise))):
            op, imm = self.leaf(e)
            code.append(encode(op, d, imm=imm))
            return
        if e.is_Add:
            self._gen_add(e, d, code)
            return
        if e.is_Mul:
            self._gen_mul(e, d, code)
            return
        if e.is_Pow:
            self._gen_pow(e, d, code)
            return
        if isinstance(e, sp.Piecewise):
            self._gen_piecewise(list(e.args), d, code)
            return
        if type(e) in _REL:
            self._gen_binary(_REL[type(e)], e.lhs, e.rhs, d, code)
            return
        if isinstance(e, (sp.And, sp.Or)):
            opn = "AND" if isinstance(e, sp.And) else "OR"
            args = sorted(e.args, key=self.need, reverse=True)
            self._gen(args[0], d, code)
            for a in args[1:]:
                self._apply_binary(opn, a, d, code)
            return
        if isinstance(e, sp.Not):
            self._gen(e.args[0], d, code)
            code.append(encode("EQ", d, d, 0, True, self._const(0.0)))
            return
        if isinstance(e, (sp.Min, sp.Max)):
            opn = "MIN" if isinstance(e, sp.Min) else "MAX"
            args = sorted(e.args, key=self.need, reverse=True)
            self._gen(args[0], d, code)
            for a in args[1:]:
                self._apply_binary(opn, a, d, code)
            return
        if e.func in _UNARY:
            self._gen(e.args[0], d, code)
            code.append(encode(_UNARY[e.func], d, d))
            return
        raise CompileError("unsupported expression in a right-hand side: %s" % (e,))

    def _apply_binary(self, opn: str, operand, d: int, code: List[int]) -> None:
        """r[d] = r[d] <op> operand."""
        if _is_number(operand):
            code.append(encode(opn, d, d, 0, True, self._const(operand)))
        else:
            self._gen(operand, d + 1, code)
            code.append(encode(opn, d, d, d + 1))

    def _gen_binary(self, opn: str, a, b, d: int, code: List[int]) -> None:
        if _is_number(a) and not _is_number(b):
            # evaluate the non-constant side first; keep operand order via a register
            self._gen(a, d, code)
            self._gen(b, d + 1, code)
            code.append(encode(opn, d, d, d + 1))
            return
        self._gen(a, d, code)
        self._apply_binary(opn, b, d, code)

    def _gen_add(self, e, d: int, code: List[int]) -> None:
        pos, neg, const = [], [], 0.0
        has_const = False
        for t in e.args:
            if _is_number(t):
                const += float(t)
                has_const = True
                continue
            c, rest = t.as_coeff_Mul()
            if _is_number(c) and float(c) < 0:
                neg.append(-t)
            else:
                pos.append(t)
        pos.sort(key=self.need, reverse=True)
        neg.sort(key=self.need, reverse=True)
        if pos:
            self._gen(pos[0], d, code)
            rest_pos = pos[1:]
        elif has_const:
            code.append(encode("CONST", d, imm=self._const(const)))
            has_const = False
            rest_pos = []
        else:
            self._gen(neg[0], d, code)
            code.append(encode("NEG", d, d))
            neg = neg[1:]
            rest_pos = []
        for t in rest_pos:
            self._apply_binary("ADD", t, d, code)
        for t in neg:
            self._apply_binary("SUB", t, d, code)
        if has_const and const != 0.0:
            code.append(encode("ADD", d, d, 0, True, self._const(const)))

    def _gen_mul(self, e, d: int, code: List[int]) -> None:
        coeff = 1.0
        num, den = [], []
        for f in e.args:
            if _is_number(f):
                coeff *= float(f)
                continue
            if f.is_Pow and _is_number(f.exp) and float(f.exp) < 0:
                ex = -f.exp
                den.append(f.base if ex == 1 else sp.Pow(f.base, ex, evaluate=False))
            else:
                num.append(f)
        num.sort(key=self.need, reverse=True)
        if num:
            self._gen(num[0], d, code)
            for f in num[1:]:
                self._apply_binary("MUL", f, d, code)
            if coeff == -1.0:
                code.append(encode("NEG", d, d))
            elif coeff != 1.0:
                code.append(encode("MUL", d, d, 0, True, self._const(coeff)))
        else:
            code.append(encode("CONST", d, imm=self._const(coeff)))
        if den:
            den.sort(key=self.need, reverse=True)
            if len(den) == 1:
                self._apply_binary("DIV", den[0], d, code)
            else:
                self._gen(sp.Mul(*den, evaluate=False), d + 1, code)
                code.append(encode("DIV", d, d, d + 1))

    def _gen_pow(self, e, d: int, code: List[int]) -> None:
        base, ex = e.base, e.exp
        if _is_number(ex):
            exf = float(ex)
            if exf == int(exf) and abs(exf) <= 1000:
                self._gen(base, d, code)
                k = int(exf)
                if k == -1:
                    code.append(encode("INV", d, d))
                elif k == 2:
                    code.append(encode("SQ", d, d))
                else:
                    code.append(encode("POWI", d, d, imm=k))
                return
            if exf == 0.5:
                self._gen(base, d, code)
                code.append(encode("SQRT", d, d))
                return
            if exf == -0.5:
                self._gen(base, d, code)
                code.append(encode("SQRT", d, d))
                code.append(encode("INV", d, d))
                return
            if base == sp.E:
                code.append(encode("CONST", d, imm=self._const(ex)))
                code.append(encode("EXP", d, d))
                return
            self._gen(base, d, code)
            code.append(encode("POW", d, d, 0, True, self._const(exf)))
            return
        if base == sp.E:
            self._gen(ex, d, code)
            code.append(encode("EXP", d, d))
            return
        self._gen_binary("POW", base, ex, d, code)

    def _gen_piecewise(self, pieces, d: int, code: List[int]) -> None:
        val, cond = pieces[0]
        if cond is sp.true or len(pieces) == 1:
            self._gen(val, d, code)
            return
        self._gen(cond, d, code)
        self._gen(val, d + 1, code)
        rest = pieces[1:]
        if len(rest) == 1 and rest[0][1] is sp.true and _is_number(rest[0][0]):
            code.append(encode("SELECT", d, d + 1, 0, True, self._const(rest[0][0])))
            return
        self._gen_piecewise(rest, d + 2, code)
        code.append(encode("SELECT", d, d + 1, d + 2))
