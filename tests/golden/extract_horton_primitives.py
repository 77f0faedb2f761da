"""Extract the HORTON primitive ERI known answers held in the reference's test
suite (tests/test_two_elec_int.py:1111-1365, ``test_two_int_brute``) into
``horton_primitive_eris.json``.

The reference file is read as text; nothing from it is imported or executed.
Each record holds the Cartesian exponents of the four centres, the exponents
(alpha, beta, gamma, delta) and the expected normalised primitive integral.
Geometry (fixed by the test's defaults, :62-66): A=(0.2,0.4,0.6),
B=(1.0,1.5,2.0), C=(0.1,0.3,0.5), D=(1.1,1.6,2.1).
"""
import ast
import json
import pathlib
import re

SRC = pathlib.Path("/root/reference/tests/test_two_elec_int.py")
OUT = pathlib.Path(__file__).with_name("horton_primitive_eris.json")


def main():
    text = SRC.read_text()
    start = text.index("def test_two_int_brute():")
    end = text.index("def test_compute_two_elec_integrals_angmom_zero_prim")
    body = text[start:end]
    calls = re.findall(r"np\.allclose\(\s*two_int_brute\((.*?)\),\s*([-0-9.e]+),?\s*\)", body,
                       flags=re.S)
    records = []
    for args, value in calls:
        node = ast.parse(f"f({args})", mode="eval").body
        pos = [ast.literal_eval(a) for a in node.args]
        kw = {k.arg: ast.literal_eval(k.value) for k in node.keywords}
        exps = dict(alpha=0.1, beta=0.2, gamma=0.3, delta=0.4)
        for name, v in zip(("alpha", "beta", "gamma", "delta"), pos[13:17]):
            exps[name] = v
        for name in ("alpha", "beta", "gamma", "delta"):
            if name in kw:
                exps[name] = kw[name]
        records.append({"a": pos[0:3], "b": pos[3:6], "c": pos[6:9], "d": pos[9:12],
                        "exps": [exps["alpha"], exps["beta"], exps["gamma"], exps["delta"]],
                        "value": float(value)})
    payload = {"source": "tests/test_two_elec_int.py:1111-1365 (HORTON values)",
               "coords": [[0.2, 0.4, 0.6], [1.0, 1.5, 2.0], [0.1, 0.3, 0.5], [1.1, 1.6, 2.1]],
               "records": records}
    OUT.write_text(json.dumps(payload, indent=1))
    print(f"wrote {len(records)} records to {OUT}")


if __name__ == "__main__":
    main()
