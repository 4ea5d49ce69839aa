"""Extract the reference's own test fixtures, as data, into reference_fixtures.json.

Reads (as text, never runs) these files of the reference checkout:
  rust/crates/gpu/src/test_util.rs   many_positions (:112-631),
                                     test_position_gradients_simple (:632-676),
                                     specific_positions_and_collider_bits (:678-2730)
  rust/crates/util/src/elastic.rs    test_lame_parameters, test_inviscid_parameters,
                                     test_viscosity_parameters (:679-696)
  rust/crates/gpu/src/radix_sort/test.rs  the keys of test_simple and specific (:13, :60)
Run once where /root/reference exists; the JSON it writes is committed.
"""
import json
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "reference_fixtures.json"

num = r"-?\d+(?:\.\d+)?(?:e[-+]?\d+)?"


def body_of(text: str, fn: str) -> str:
    start = text.index(f"pub fn {fn}(")
    depth = 0
    i = text.index("{", start)
    for j in range(i, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[i : j + 1]
    raise ValueError(fn)


tu = (REF / "rust/crates/gpu/src/test_util.rs").read_text()
many = [[float(a), float(b), float(c)] for a, b, c in re.findall(rf"Vector3::new\(({num}),\s*({num}),\s*({num})\)", body_of(tu, "many_positions"))]
simple_body = body_of(tu, "test_position_gradients_simple")
simple = []
if "Matrix3::identity()" in simple_body:
    simple.append([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
for block in re.findall(r"from_row_slice\(&\[(.*?)\]\)", simple_body, re.S):
    vals = [float(v) for v in re.findall(num, block)]
    simple.append([vals[0:3], vals[3:6], vals[6:9]])
spec_body = body_of(tu, "specific_positions_and_collider_bits")
specific = [
    {"position": [float(a), float(b), float(c)], "collider_bits": int(bits)}
    for a, b, c, bits in re.findall(
        rf"position:\s*Vector3::new\(({num}),\s*({num}),\s*({num})\),\s*collider_bits:\s*(\d+)", spec_body
    )
]

el = (REF / "rust/crates/util/src/elastic.rs").read_text()
lame_block = re.search(r"fn test_lame_parameters.*?\[(\[.*?\])\]\s*\.into_iter", el, re.S).group(1)
lame = [[float(a), float(b)] for a, b in re.findall(rf"\[({num})\.?,\s*({num})\.?\]", lame_block)]
inv_block = re.search(r"fn test_inviscid_parameters.*?\[(\(.*?\))\]\s*\.into_iter", el, re.S).group(1)
inviscid = [[float(a), int(b)] for a, b in re.findall(rf"\(({num})\.?,\s*(\d+)\)", inv_block)]

# test_viscosity_parameters: (a..b).flat_map(|i| (c..d).map(|j| [i, j])), the
# (dynamic, bulk) pairs of a product of two integer ranges
visc_body = body_of(el, "test_viscosity_parameters")
(i0, i1), (j0, j1) = [(int(a), int(b)) for a, b in re.findall(r"\((\d+)\.\.(\d+)\)", visc_body)]
viscosity = [[float(i), float(j)] for i in range(i0, i1) for j in range(j0, j1)]

rs = (REF / "rust/crates/gpu/src/radix_sort/test.rs").read_text()
sort_simple = [int(k) for k in re.search(r"let keys = \[([\d,\s]+)\];", rs).group(1).split(",") if k.strip()]
sort_specific = [int(k) for k in re.findall(r"let keys = \[([\d,\s]+)\];", rs)[1].split(",") if k.strip()]

data = {
    "source": "Algebraic-UG/squishy_volumes test fixtures, extracted as data by make_reference_fixtures.py",
    "many_positions": many,
    "position_gradients_simple_row_major": simple,
    "specific_positions_and_collider_bits": specific,
    "lame_parameters_E_nu": lame,
    "inviscid_parameters_K_exponent": inviscid,
    "viscosity_parameters_dynamic_bulk": viscosity,
    "radix_sort_simple_keys": sort_simple,
    "radix_sort_specific_keys": sort_specific,
}
assert len(many) == 512, len(many)
assert len(simple) == 9, len(simple)
assert len(specific) == 512, len(specific)
assert len(lame) == 4 and len(inviscid) == 4 and len(viscosity) == 64
OUT.write_text(json.dumps(data))
print(f"wrote {OUT}: {len(many)} positions, {len(simple)} F, {len(specific)} specific, lame {lame}, inviscid {inviscid}, sort {sort_simple} / {sort_specific}")
