"""Kernel expressions of ferebus.toml -> postfix kernel programs for the C ABI.

Mirrors the reference's interpreter (src/interpreter/):

* lexer: identifiers, numbers, ``+ - * / ( )`` (src/interpreter/lexer.f90);
* grammar (src/interpreter/parser.f90:102-197)::

      expr   := term (('+' | '-') term)*          left associative
      term   := factor (('*' | '/') term)?        right recursive
      factor := '+' factor | '-' factor | NUMBER | '(' expr ')' | IDENT

* visitor (src/interpreter/ast.f90:173-271): ``+`` -> sum kernel, ``*`` ->
  product kernel, identifiers -> the kernel configured under
  ``[kernels.<name>]``, numbers -> a constant kernel whose value is one PSO
  parameter.  ``-`` and ``/`` (binary or unary) are rejected, as the reference
  stops with "not yet implemented".

Postfix order lists leaves left to right, which is the order the reference
stacks kernel parameters in theta (src/kernels/kernel_module.f90:227-245).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

LEAF_RBF = 1
LEAF_PERIODIC = 2
LEAF_RBF_CYCLIC = 3
LEAF_LINEAR = 4
LEAF_CONSTANT = 5
OP_SUM = 16
OP_PROD = 17

# [kernels.<name>] type strings (src/config/config_module.f90:171-183)
KERNEL_TYPES = {
    "rbf": LEAF_RBF,
    "rbf-cyclic": LEAF_RBF_CYCLIC,
    "periodic": LEAF_PERIODIC,
    "linear": LEAF_LINEAR,
}

# lexer.f90:95-160: numbers start with a digit (optional fraction and
# exponent); identifiers are a letter followed by letters and digits
_TOKEN = re.compile(r"\s*(?:(?P<num>\d+(?:\.\d*)?(?:[eE][-+]?\d*)?)"
                    r"|(?P<id>[A-Za-z][A-Za-z0-9]*)|(?P<op>[-+*/()]))")


class KernelExpressionError(ValueError):
    """The kernel expression cannot be built (the reference stops here)."""


@dataclass
class KernelConfig:
    """One ``[kernels.<name>]`` table."""

    name: str
    type: str = "rbf"
    active_dims: Optional[Sequence[int]] = None   # 1-based; None = all features


@dataclass
class KernelProgram:
    """Postfix program in the layout of ``fb_kernel_spec``."""

    node_type: List[int] = field(default_factory=list)
    node_ndim: List[int] = field(default_factory=list)
    active_dims: List[int] = field(default_factory=list)
    leaf_names: List[str] = field(default_factory=list)
    composition: str = ""

    @property
    def n_params(self) -> int:
        """Kernel parameters (the nugget, when optimised, comes on top)."""
        n = 0
        for t, nd in zip(self.node_type, self.node_ndim):
            if t == LEAF_CONSTANT:
                n += 1
            elif t < OP_SUM:
                n += nd
        return n


def _tokens(text: str) -> List[tuple]:
    out, pos = [], 0
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise KernelExpressionError(f"unexpected character in kernel expression at {text[pos:]!r}")
        pos = m.end()
        if m.group("num"):
            out.append(("num", float(m.group("num"))))
        elif m.group("id"):
            out.append(("id", m.group("id")))
        else:
            out.append(("op", m.group("op")))
    out.append(("eof", None))
    return out


class _Parser:
    def __init__(self, text: str):
        self.toks = _tokens(text)
        self.i = 0

    def peek(self):
        return self.toks[self.i]

    def eat(self, kind, value=None):
        tok = self.toks[self.i]
        if tok[0] != kind or (value is not None and tok[1] != value):
            raise KernelExpressionError(f"expected {value or kind}, found {tok[1]!r}")
        self.i += 1
        return tok

    def factor(self):
        tok = self.peek()
        if tok == ("op", "+"):
            self.eat("op", "+")
            return ("unary", "+", self.factor())
        if tok == ("op", "-"):
            self.eat("op", "-")
            return ("unary", "-", self.factor())
        if tok[0] == "num":
            self.eat("num")
            return ("num", tok[1])
        if tok == ("op", "("):
            self.eat("op", "(")
            node = self.expr()
            self.eat("op", ")")
            return node
        if tok[0] == "id":
            self.eat("id")
            return ("var", tok[1])
        raise KernelExpressionError(f"unexpected token {tok[1]!r}")

    def term(self):
        node = self.factor()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.eat("op")[1]
            node = ("bin", op, node, self.term())
        return node

    def expr(self):
        node = self.term()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.eat("op")[1]
            node = ("bin", op, node, self.term())
        return node

    def parse(self):
        node = self.expr()
        if self.peek()[0] != "eof":
            raise KernelExpressionError(f"trailing input {self.peek()[1]!r} in kernel expression")
        return node


def parse(text: str):
    """Abstract syntax tree of a kernel expression."""
    return _Parser(text).parse()


def compile_kernel(text: str, kernels: Dict[str, KernelConfig], nfeats: int) -> KernelProgram:
    """Postfix kernel program for ``text`` over ``nfeats`` features."""
    prog = KernelProgram()
    names: List[str] = []

    def visit(node) -> str:
        kind = node[0]
        if kind == "bin":
            op = node[1]
            if op in ("-", "/"):
                raise KernelExpressionError(f"Kernel {'minus' if op == '-' else 'divide'} not yet implemented")
            left = visit(node[2])
            right = visit(node[3])
            prog.node_type.append(OP_SUM if op == "+" else OP_PROD)
            prog.node_ndim.append(0)
            return f"({left}{op}{right})"
        if kind == "unary":
            if node[1] == "-":
                raise KernelExpressionError("Kernel minus not yet implemented")
            return visit(node[2])
        if kind == "num":
            prog.node_type.append(LEAF_CONSTANT)
            prog.node_ndim.append(0)
            prog.leaf_names.append("c")
            return "c"
        name = node[1]
        if name not in kernels:
            raise KernelExpressionError(f"kernel {name!r} is not configured under [kernels]")
        cfg = kernels[name]
        ktype = cfg.type.lower()
        if ktype not in KERNEL_TYPES:
            raise KernelExpressionError(f"unknown kernel type {cfg.type!r}")
        # kernel_type_config.f90:54-73: default active dims are 1..nfeats
        dims = list(cfg.active_dims) if cfg.active_dims is not None else list(range(1, nfeats + 1))
        if not dims or any(not (1 <= int(f) <= nfeats) for f in dims):
            raise KernelExpressionError(f"kernel {name!r}: active_dimensions outside 1..{nfeats}")
        if KERNEL_TYPES[ktype] == LEAF_RBF_CYCLIC and dims != list(range(1, nfeats + 1)):
            raise KernelExpressionError("rbf-cyclic sums over every feature; active_dimensions must be 1..nfeats")
        prog.node_type.append(KERNEL_TYPES[ktype])
        prog.node_ndim.append(len(dims))
        prog.active_dims.extend(int(f) for f in dims)
        prog.leaf_names.append(name)
        names.append(name)
        return name

    prog.composition = visit(parse(text))
    return prog
