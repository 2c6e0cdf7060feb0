"""Write tests/golden/bruss_dx025.json from the reference's generated Brusselator RHS.

Source (read as text, run once in the build container):
  /root/reference/docs/src/generated/bruss_code.md — the in-place function that
  ModelingToolkit generated for the Brusselator at dx = dy = 1/4 (32 unknowns).

The fixture is data only: the unknown ordering, the constant / linear / cubic
coefficients of each of the 32 right-hand-side expressions, and the values of the
expressions on seeded random states.  Nothing of the generated code's text is stored.

Usage: python tests/golden/make_bruss_fixture.py
"""
import json
import os
import re
import sys

import numpy as np

SRC = "/root/reference/docs/src/generated/bruss_code.md"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bruss_dx025.json")


def tokenize(s):
    toks = re.findall(r'\(\+\)|\(\*\)|\(\^\)|var"[^"]+"|-?\d+\.\d+(?:e-?\d+)?|-?\d+|[(),]', s)
    return toks


def parse(toks, i=0):
    t = toks[i]
    if t in ("(+)", "(*)", "(^)"):
        op = t[1]
        assert toks[i + 1] == "("
        i += 2
        args = []
        while True:
            a, i = parse(toks, i)
            args.append(a)
            if toks[i] == ",":
                i += 1
                continue
            assert toks[i] == ")"
            i += 1
            break
        return (op, args), i
    if t.startswith('var"'):
        return ("var", t[4:-1]), i + 1
    return ("num", float(t)), i + 1


def evaluate(node, env):
    kind = node[0]
    if kind == "num":
        return node[1]
    if kind == "var":
        return env[node[1]]
    op, args = node
    vals = [evaluate(a, env) for a in args]
    if op == "+":
        out = vals[0]
        for v in vals[1:]:
            out = out + v
        return out
    if op == "*":
        out = vals[0]
        for v in vals[1:]:
            out = out * v
        return out
    if op == "^":
        return vals[0] ** vals[1]
    raise ValueError(op)


def terms(node, sign=1.0):
    """Flatten a sum into (coefficient, monomial) terms; monomial = sorted tuple of var names."""
    if node[0] == "+":
        out = []
        for a in node[1]:
            out += terms(a, sign)
        return out
    if node[0] == "num":
        return [(sign * node[1], ())]
    if node[0] == "var":
        return [(sign, (node[1],))]
    if node[0] == "*":
        coef, mono = sign, []
        for a in node[1]:
            for c, m in terms(a):
                coef *= c
                mono += list(m)
        return [(coef, tuple(sorted(mono)))]
    if node[0] == "^":
        base, ex = node[1]
        assert base[0] == "var" and ex[0] == "num"
        return [(sign, tuple([base[1]] * int(ex[1])))]
    raise ValueError(node)


def main():
    text = open(SRC).read()
    # the in-place body: lines "ˍ₋out[k] = <expr>"
    lines = re.findall(r"ˍ₋out\[(\d+)\] = (.*)", text)
    names = re.findall(r'var"([uv]\[\d, \d\]\(t\))" = @inbounds\(ˍ₋arg1\[(\d+)\]\)', text)
    order = [None] * 32
    for nm, k in names[:32]:
        order[int(k) - 1] = nm
    exprs = {}
    for k, e in lines:
        e = re.sub(r"\[[^\]]*\]\(https?://[^)]*\)", "", e)  # drop any markdown link noise
        toks = tokenize(e)
        tree, i = parse(toks)
        assert i == len(toks), "unparsed tail in output %s" % k
        exprs[int(k) - 1] = tree
    assert len(exprs) == 32 and all(o is not None for o in order)
    rng = np.random.default_rng(20240611)
    samples = []
    for s in range(6):
        x = rng.uniform(0.5, 1.5, 32) if s < 5 else np.arange(1, 33) / 7.0
        env = {order[i]: float(x[i]) for i in range(32)}
        y = [evaluate(exprs[k], env) for k in range(32)]
        samples.append({"u": x.tolist(), "du": y})
    structure = []
    for k in range(32):
        tl = terms(exprs[k])
        const = sum(c for c, m in tl if len(m) == 0)
        lin = {}
        cubic = []
        for c, m in tl:
            if len(m) == 1:
                lin[order.index(m[0])] = lin.get(order.index(m[0]), 0.0) + c
            elif len(m) == 3:
                cubic.append([c, [order.index(v) for v in m]])
        structure.append({"const": const, "linear": sorted([[i, c] for i, c in lin.items()]), "cubic": cubic})
    json.dump({"source": "docs/src/generated/bruss_code.md (dx = dy = 1/4)", "unknowns": order,
               "samples": samples, "structure": structure}, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
