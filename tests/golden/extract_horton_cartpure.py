"""Extract the HORTON Cartesian->pure matrices held in the reference's test suite
(tests/test_spherical.py:191-505: ``test_generate_transformation_horton*``) into
``horton_cartpure.json``.

The reference file is read as text; its numeric expressions (such as
``-1 / 4 * 6**0.5``) are evaluated by a restricted arithmetic evaluator, and
nothing from the reference is imported or executed.  Each record holds
``l``, the Cartesian order, the pure-function order and the dense matrix
(shape (2l+1, n_cart), "left" orientation).
"""
import ast
import itertools
import json
import operator
import pathlib

SRC = pathlib.Path("/root/reference/tests/test_spherical.py")
OUT = pathlib.Path(__file__).with_name("horton_cartpure.json")
_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow}


def _num(node):
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_num(node.operand)
    if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
        return _OPS[type(node.op)](_num(node.left), _num(node.right))
    if isinstance(node, (ast.List, ast.Tuple)):
        return [_num(e) for e in node.elts]
    if isinstance(node, ast.Call) and getattr(node.func, "attr", None) == "array":
        return _num(node.args[0])
    if isinstance(node, ast.ListComp):
        # HORTON Cartesian order written as a comprehension over
        # it.combinations_with_replacement(range(3), l): component counts.
        gen = node.generators[0].iter
        if getattr(gen.func, "attr", "") == "combinations_with_replacement":
            l = _num(gen.args[1])
            return [[c.count(0), c.count(1), c.count(2)]
                    for c in itertools.combinations_with_replacement(range(3), l)]
    raise ValueError(ast.dump(node))


def main():
    tree = ast.parse(SRC.read_text())
    records = []
    for fn in tree.body:
        if not (isinstance(fn, ast.FunctionDef) and fn.name.startswith(
                "test_generate_transformation_horton")):
            continue
        answer = None
        triples = None
        for top in fn.body:
            if isinstance(top, ast.Assign) and isinstance(top.targets[0], ast.Name):
                name = top.targets[0].id
                if name == "answer" and getattr(top.value.func, "attr", "") == "array":
                    answer = _num(top.value)
                    triples = None
                elif name == "horton_transform":
                    triples = _num(top.value)
                    answer = None
                continue
            for stmt in ast.walk(top):
                if not (isinstance(stmt, ast.Call) and getattr(stmt.func, "id", "") ==
                        "generate_transformation"):
                    continue
                l = _num(stmt.args[0])
                cart = _num(stmt.args[1])
                sph = [e.value for e in stmt.args[2].elts]
                if triples is not None:
                    mat = [[0.0] * len(cart) for _ in range(2 * l + 1)]
                    for r, c, v in triples:
                        # the test negates the rows it labels "-c{m}"/"-s{m}"
                        # (test_spherical.py:334-339)
                        mat[r][c] = -v if sph[r].startswith("-") else v
                else:
                    mat = answer
                records.append({"l": l, "cart": cart, "sph": sph, "matrix": mat})
    OUT.write_text(json.dumps({"source": "tests/test_spherical.py:191-505 (HORTON cartpure)",
                               "records": records}, indent=1))
    print(f"wrote {len(records)} records to {OUT}")


if __name__ == "__main__":
    main()
