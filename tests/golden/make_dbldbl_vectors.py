"""Writes tests/golden/dbldblarith_kat.json: the known-answer vectors of the reference's
arithmetic unit test (tests/src/misc/dbldblarith.c:36-45 inputs, :71-88 sum and :52-69
product expectations), transcribed as data.  The values are decimal literals from that
file; the script only normalises them to exact hex floats.
"""
import json
import os

a = 0.000000034567          # dbldblarith.c:37
b = 2.34567                 # dbldblarith.c:40
c = 12345678.9              # dbldblarith.c:43
vectors = {
    "inputs": {"a": a.hex(), "b": b.hex(), "c": c.hex()},
    # ((c + b) + a) + a in double-double: SCIPdbldblSum then 2x SCIPdbldblSum21 (dbldblarith.c:79-81)
    "sum": {"expected": (1.2345681245670069507e7).hex(), "naive_double_differs": True},
    # (a * b) * c: SCIPdbldblProd then SCIPdbldblProd21 (dbldblarith.c:58-59)
    "product": {"expected": (1.0010219031129229441).hex(), "naive_double_differs": True},
}
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dbldblarith_kat.json")
with open(out, "w") as f:
    json.dump(vectors, f, indent=1)
print("wrote", out)
