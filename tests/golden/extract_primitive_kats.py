"""Extract the closed-form known-answer tables of the reference's primitive collision tests.

Reads ``newton/tests/test_collision_primitives.py`` of the reference checkout as text (no
import), evaluates each ``test_cases`` literal of the listed test methods together with the
simple local assignments before it (``wp.mat33(...)`` becomes a row-major tuple), and writes
``primitive_kats.json`` next to this script: per test, the input rows and the expected
distance (and contact count where the table has one).

Usage: python tests/golden/extract_primitive_kats.py /root/reference
"""

import ast
import json
import os
import sys

TESTS = {
    "test_plane_sphere": "plane_sphere",
    "test_sphere_sphere": "sphere_sphere",
    "test_sphere_capsule": "sphere_capsule",
    "test_capsule_capsule": "capsule_capsule",
    "test_sphere_box": "sphere_box",
    "test_plane_capsule": "plane_capsule",
    "test_plane_box": "plane_box",
    "test_capsule_box": "capsule_box",
    "test_box_box": "box_box",
}


_WP_CALLS = ("mat33", "vec3")  # wp.mat33(...) / wp.vec3(...) become row-major tuples


def _eval(node, env):
    """Value of a literal expression: numbers, tuples / lists, unary minus / plus, names bound
    earlier in the same test, and wp.mat33 / wp.vec3 calls on such values.  Anything else raises
    ValueError; nothing from the file is executed."""
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float, bool)):
        return node.value
    if isinstance(node, (ast.Tuple, ast.List)):
        vals = [_eval(e, env) for e in node.elts]
        return tuple(vals) if isinstance(node, ast.Tuple) else vals
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _eval(node.operand, env)
        if not isinstance(v, (int, float)):
            raise ValueError("unary sign on a non-number")
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.Name) and node.id in env:
        return env[node.id]
    if (isinstance(node, ast.Call) and not node.keywords and isinstance(node.func, ast.Attribute)
            and isinstance(node.func.value, ast.Name) and node.func.value.id == "wp"
            and node.func.attr in _WP_CALLS):
        args = [_eval(a, env) for a in node.args]
        if not all(isinstance(a, (int, float)) for a in args):
            raise ValueError("wp constructor with non-numeric arguments")
        return tuple(float(a) for a in args)
    raise ValueError(f"unsupported expression {type(node).__name__}")


def extract(ref_root: str) -> dict:
    path = os.path.join(ref_root, "newton", "tests", "test_collision_primitives.py")
    tree = ast.parse(open(path).read())
    out = {}
    for node in ast.walk(tree):
        if not isinstance(node, ast.FunctionDef) or node.name not in TESTS:
            continue
        env = {}
        for stmt in node.body:
            if not isinstance(stmt, ast.Assign) or len(stmt.targets) != 1:
                continue
            tgt = stmt.targets[0]
            if not isinstance(tgt, ast.Name):
                continue
            try:
                val = _eval(stmt.value, env)
            except ValueError:
                continue
            env[tgt.id] = val
            if tgt.id == "test_cases":
                break
        if "test_cases" in env:
            out[TESTS[node.name]] = {"source": f"newton/tests/test_collision_primitives.py::{node.name}",
                                     "cases": json.loads(json.dumps(env["test_cases"]))}
    return out


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = extract(ref)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "primitive_kats.json")
    with open(dst, "w") as f:
        json.dump(data, f)
    print(f"wrote {dst}: " + ", ".join(f"{k}({len(v['cases'])})" for k, v in data.items()))
