"""Write tests/golden/serde_order.json: the declaration order of the fields of
every serde struct, and of the variants of every serde enum, in the
reference's Rust sources (src/**/*.rs, read as text).

bincode writes a struct as its fields in declaration order and an enum as
its variant's index, so this order is the layout of the reference's binary
recovery snapshot (src/integrator/output.rs:28-54).  The fixture holds names
only; tests/test_bincode.py checks the snapshot codec's schema against it.

Run from the repository root: ``python tests/golden/make_serde_order.py``.
No reference code is imported or executed.
"""
import glob
import json
import os
import re

SRC = "/root/reference/src"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "serde_order.json")


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def blocks(text, kind):
    """(name, body) of every `pub <kind> Name { ... }` preceded by a derive
    that includes Serialize."""
    out = []
    for m in re.finditer(r"#\[derive\(([^)]*)\)\]\s*pub\s+" + kind + r"\s+(\w+)\s*\{", text):
        if "Serialize" not in m.group(1):
            continue
        depth, i = 1, m.end()
        while depth:
            depth += {"{": 1, "}": -1}.get(text[i], 0)
            i += 1
        out.append((m.group(2), text[m.end():i - 1]))
    return out


def main():
    structs, enums = {}, {}
    for path in sorted(glob.glob(os.path.join(SRC, "**", "*.rs"), recursive=True)):
        text = strip_comments(open(path).read())
        for name, body in blocks(text, "struct"):
            body = re.sub(r"#\[[^\]]*\]", "", body)
            structs[name] = re.findall(r"(?:pub\s+)?(\w+)\s*:", body)
        for name, body in blocks(text, "enum"):
            variants, depth, cur = [], 0, ""
            for ch in body:
                if ch in "({":
                    depth += 1
                elif ch in ")}":
                    depth -= 1
                if ch == "," and depth == 0:
                    variants.append(cur)
                    cur = ""
                else:
                    cur += ch
            variants.append(cur)
            enums[name] = [re.match(r"\s*(\w+)", v).group(1) for v in variants if v.strip()]
    with open(OUT, "w") as f:
        json.dump({"structs": structs, "enums": enums}, f, indent=1, sort_keys=True)
    print(f"{len(structs)} structs, {len(enums)} enums -> {OUT}")


if __name__ == "__main__":
    main()
