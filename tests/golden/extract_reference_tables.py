"""Writes tests/golden/reference_tables.json: the model tables of finitewave's
16 ionic model classes, read from the reference's source text with `ast`
(parsed, never imported or executed).

For each class of cpuwave2D/model/*.py and cpuwave3D/model/*.py:
  * the attributes `__init__` assigns (`self.X = <number or arithmetic of
    numbers and earlier self attributes>`), D_model, state_vars and npfloat,
    with the parent class's table merged in for the 3D classes that only call
    `super().__init__()`;
  * `initialize`: which `init_*` attribute each state array is filled from
    (`self.X = self.init_Y * np.ones...`) and whether u_new is a copy of u;
  * `run_ionic_kernel`: the kernel called and its arguments as written, and
    the kernel function's own parameter names, so the argument slot of every
    array and scalar is known (Courtemanche passes self.xr into the slot
    named xs, courtemanche_2d.py:149 vs :569).

Run in a container that holds the reference source:
    python tests/golden/extract_reference_tables.py /root/reference
"""
from __future__ import annotations

import ast
import json
import operator
import sys
from pathlib import Path

OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
       ast.Div: operator.truediv, ast.Pow: operator.pow}


def _self_attr(node):
    if (isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name)
            and node.value.id == "self"):
        return node.attr
    return None


def _value(node, table):
    """A number from literals, + - * / ** and attributes already in table;
    strings and lists of strings as they are; None for anything else."""
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        v = _value(node.operand, table)
        return -v if isinstance(v, (int, float)) else None
    if isinstance(node, ast.BinOp) and type(node.op) in OPS:
        a, b = _value(node.left, table), _value(node.right, table)
        if isinstance(a, (int, float)) and isinstance(b, (int, float)):
            return OPS[type(node.op)](a, b)
        return None
    name = _self_attr(node)
    if name is not None:
        return table.get(name)
    if isinstance(node, ast.List):
        items = [_value(e, table) for e in node.elts]
        return items if all(isinstance(i, str) for i in items) else None
    return None


def _method(cls, name):
    for item in cls.body:
        if isinstance(item, ast.FunctionDef) and item.name == name:
            return item
    return None


def _init_table(cls):
    table = {}
    fn = _method(cls, "__init__")
    if fn is None:
        return table
    for st in ast.walk(fn):
        if isinstance(st, ast.Assign) and len(st.targets) == 1:
            name = _self_attr(st.targets[0])
            if name is None:
                continue
            v = _value(st.value, table)
            if v is not None:
                table[name] = v
    return table


def _init_sources(cls):
    """{state attribute: init attribute} from initialize(), and whether
    u_new = u.copy()."""
    fn = _method(cls, "initialize")
    out, copy_u_new = {}, False
    if fn is None:
        return out, None
    for st in ast.walk(fn):
        if not (isinstance(st, ast.Assign) and len(st.targets) == 1):
            continue
        name = _self_attr(st.targets[0])
        if name is None:
            continue
        v = st.value
        if isinstance(v, ast.BinOp) and isinstance(v.op, ast.Mult):
            src = _self_attr(v.left)
            if src is not None and src.startswith("init_"):
                out[name] = src
        if (name == "u_new" and isinstance(v, ast.Call) and isinstance(v.func, ast.Attribute)
                and v.func.attr == "copy" and _self_attr(v.func.value) == "u"):
            copy_u_new = True
    return out, copy_u_new


def _kernel_call(cls):
    fn = _method(cls, "run_ionic_kernel")
    if fn is None:
        return None
    for st in ast.walk(fn):
        if isinstance(st, ast.Call) and isinstance(st.func, ast.Name):
            return st.func.id, [ast.unparse(a) for a in st.args]
    return None


def _functions(tree):
    return {n.name: [a.arg for a in n.args.args] for n in tree.body
            if isinstance(n, ast.FunctionDef)}


def extract(root: Path) -> dict:
    pkg = root / "finitewave"
    trees, classes = {}, {}
    for sub in ("cpuwave2D", "cpuwave3D"):
        for path in sorted((pkg / sub / "model").glob("*.py")):
            if path.name == "__init__.py":
                continue
            tree = ast.parse(path.read_text())
            rel = str(path.relative_to(pkg))
            trees[rel] = tree
            for node in tree.body:
                if isinstance(node, ast.ClassDef):
                    classes[node.name] = (rel, node)
    # kernel parameter names by (file, function), also through imports
    funcs = {rel: _functions(t) for rel, t in trees.items()}
    out = {}
    for name, (rel, cls) in sorted(classes.items()):
        bases = [b.id for b in cls.bases if isinstance(b, ast.Name)]
        table = {}
        init_src, copy_u_new = {}, None
        for b in bases:  # the 3D classes extend their 2D class
            if b in classes:
                table.update(_init_table(classes[b][1]))
                s, c = _init_sources(classes[b][1])
                init_src.update(s)
                copy_u_new = c if copy_u_new is None else copy_u_new
        table.update(_init_table(cls))
        s, c = _init_sources(cls)
        init_src.update(s)
        if c is not None and _method(cls, "initialize") is not None:
            copy_u_new = c
        call = _kernel_call(cls)
        kname, args = call
        params = funcs[rel].get(kname)
        if params is None:  # defined in a module the class file imports
            for other in funcs.values():
                if kname in other:
                    params = other[kname]
                    break
        idx = args.index("self.cardiac_tissue.myo_indexes")
        state_args = [a.split(".", 1)[1] for a in args[2:idx]]
        scalar_args = [a.split(".", 1)[1] for a in args[idx + 2:]]
        out[name] = dict(
            file=rel, bases=bases, kernel=kname,
            kernel_params=params,
            state_args=state_args,                    # attributes, in the call's order
            state_slots=params[2:idx],                # the kernel's names for those slots
            scalar_args=scalar_args,                  # attributes after dt
            scalars=[table[a] for a in scalar_args],  # their default values
            init=[table[init_src[a]] for a in state_args],
            init_u=table.get("init_u"),
            D_model=table.get("D_model"),
            state_vars=table.get("state_vars"),
            npfloat=table.get("npfloat"),
            copy_u_new=bool(copy_u_new),
        )
    return out


def main(argv) -> int:
    root = Path(argv[1] if len(argv) > 1 else "/root/reference")
    data = extract(root)
    dest = Path(__file__).resolve().parent / "reference_tables.json"
    dest.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
    print(f"{dest}: {len(data)} classes")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
